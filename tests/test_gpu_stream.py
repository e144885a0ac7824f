"""Persistent mixed-stream decoder (cusat_stream_*) on the GPU.

A stream of frames of several MODCODs (normal and short FECFRAMEs,
interleaved, LLRs packed back to back) decoded in one launch must give, frame
for frame, exactly what each frame's own context gives (cusat_ldpc_decode +
cusat_bch_check; cusat_rx_fec_learned for learned codes), whatever the
tuning (6-wave or 12-wave split-edge workgroups included); a hipGraph capture of the launch replays correctly with new LLRs.
"""
import numpy as np
import pytest

import cusat_amd as ca
from tests.fec_util import snr_scales

pytestmark = pytest.mark.gpu

MIX = [(0, "1/2", "QPSK", 1.5), (0, "3/4", "8PSK", 7.9), (0, "2/3", "16APSK", 9.4),
       (0, "9/10", "8PSK", 11.3), (0, "1/4", "QPSK", 0.1), (1, "1/2", "QPSK", 2.4),
       (1, "3/5", "8PSK", 6.6), (1, "5/6", "16APSK", 11.6), (1, "1/4", "QPSK", 0.4)]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def build_stream(mix, per_code, seed, max_trials=50, contexts=None):
    """Contexts, frame records (shuffled), packed device LLRs, and the
    per-frame (context index, row in that context's batch)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    ctxs = contexts or [ca.FecContext(fs, r, m, max_trials=max_trials, max_batch=per_code) for fs, r, m, _ in mix]
    llr_by_code = []
    for k, (fs, r, m, es) in enumerate(mix):
        sym, _, _ = ca.gen_frames(fs, r, m, seed * 100 + k, per_code, es)
        nv, g = snr_scales(es)
        d = torch.empty((per_code, ctxs[k].info["n_ldpc"]), dtype=torch.int16, device="cuda")
        ctxs[k].demap(torch.from_numpy(sym).cuda(), d, per_code, nv, g)
        llr_by_code.append(d)
    who = [(k, j) for k in range(len(mix)) for j in range(per_code)]
    perm = rng.permutation(len(who))
    who = [who[p] for p in perm]
    frames = np.zeros(len(who), ca.STREAM_FRAME)
    lo = ho = 0
    parts = []
    for n, (k, j) in enumerate(who):
        frames[n] = (k, 0, lo, ho)
        parts.append(llr_by_code[k][j])
        lo += ctxs[k].info["n_ldpc"]
        ho += ctxs[k].info["k_ldpc"] // 8
    llr = torch.cat(parts).contiguous()
    torch.cuda.synchronize()
    return ctxs, frames, llr, llr_by_code, who, ho


def per_code_reference(ctxs, llr_by_code, learned_trials=None):
    torch = _torch()
    out = []
    for c, d in zip(ctxs, llr_by_code):
        B = d.shape[0]
        h = torch.zeros((B, c.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
        t = torch.full((B,), -7, dtype=torch.int32, device="cuda")
        v = torch.full((B,), -7, dtype=torch.int32, device="cuda")
        fb = torch.zeros(B, dtype=torch.int32, device="cuda")
        if learned_trials and getattr(c, "_learned", False):
            c.rx_fec_learned(d, h, t, v, fb, B, learned_trials)
        else:
            c.decode(d, h, t, B)
            c.bch_check(h, v, B)
        torch.cuda.synchronize()
        out.append((h.cpu().numpy(), t.cpu().numpy(), v.cpu().numpy(), fb.cpu().numpy()))
    return out


def run_stream(sc, frames, llr, hard_bytes):
    torch = _torch()
    n = len(frames)
    hard = torch.zeros(hard_bytes, dtype=torch.uint8, device="cuda")
    tr = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    va = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    fb = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    sc.decode(llr, hard, tr, va, fb)
    torch.cuda.synchronize()
    return hard.cpu().numpy(), tr.cpu().numpy(), va.cpu().numpy(), fb.cpu().numpy()


def check(ctxs, frames, who, ref, got):
    hard, tr, va, fb = got
    for n, (k, j) in enumerate(who):
        kb = ctxs[k].info["k_ldpc"] // 8
        h0, t0, v0, f0 = ref[k]
        o = frames[n]["hard_offset"]
        assert tr[n] == t0[j], (n, k, j)
        assert va[n] == v0[j], (n, k, j)
        assert fb[n] == f0[j], (n, k, j)
        np.testing.assert_array_equal(hard[o:o + kb], h0[j])


@pytest.mark.parametrize("tuning", [None, {"wg_per_cu": 1, "frames_per_job": 4, "order": 0, "split_edges": 0},
                                    {"wg_per_cu": 1, "frames_per_job": 2, "order": 1, "split_edges": 0},
                                    {"wg_per_cu": 1, "frames_per_job": 1, "order": 1, "split_edges": 1},
                                    {"wg_per_cu": 1, "frames_per_job": 3, "order": 0, "split_edges": 1}])
def test_stream_matches_per_code_decoders(tuning):
    ctxs, frames, llr, by_code, who, hb = build_stream(MIX, 5, 3)
    ref = per_code_reference(ctxs, by_code)
    t = np.concatenate([r[1] for r in ref])
    assert (t < 0).any() and (t >= 0).any(), "the mix needs failing and converging frames"
    sc = ca.StreamContext(ctxs, len(frames))
    sc.plan(frames, tuning)
    check(ctxs, frames, who, ref, run_stream(sc, frames, llr, hb))
    # a second launch on the same plan (the queue reset itself)
    check(ctxs, frames, who, ref, run_stream(sc, frames, llr, hb))


def test_stream_autotune_keeps_results():
    torch = _torch()
    ctxs, frames, llr, by_code, who, hb = build_stream(MIX[:4] + MIX[5:7], 6, 5)
    ref = per_code_reference(ctxs, by_code)
    sc = ca.StreamContext(ctxs, len(frames))
    sc.plan(frames)
    n = len(frames)
    hard = torch.zeros(hb, dtype=torch.uint8, device="cuda")
    tr = torch.zeros(n, dtype=torch.int32, device="cuda")
    va = torch.zeros(n, dtype=torch.int32, device="cuda")
    chosen, ms = sc.autotune(llr, hard, tr, va, reps=1)
    assert ms > 0 and chosen == {k: sc.tuning()[k] for k in chosen}
    assert chosen["split_edges"] == 1  # the 6-wave kernel is only the fallback
    check(ctxs, frames, who, ref, run_stream(sc, frames, llr, hb))


def test_stream_learned_matches_rx_fec_learned():
    from cusat_amd import learn
    _, bq, aq = learn.load_params()
    # 8.9 dB: 9 of these 12 frames fail the 20 learned iterations (checked with the oracle)
    mix = [(0, "2/3", "16APSK", 8.9), (0, "1/2", "QPSK", 2.1), (1, "1/2", "QPSK", 2.4)]
    ctxs = [ca.FecContext(fs, r, m, max_trials=50, max_batch=12) for fs, r, m, _ in mix]
    ctxs[0].set_params(bq, aq)
    ctxs[0]._learned = True
    ctxs, frames, llr, by_code, who, hb = build_stream(mix, 12, 7, contexts=ctxs)
    ref = per_code_reference(ctxs, by_code, learned_trials=bq.shape[0])
    assert ref[0][3].any() and not ref[0][3].all(), "the case needs frames with and without fallback"
    sc = ca.StreamContext(ctxs, len(frames), ca.STREAM_LEARNED)
    for split in (0, 1):
        sc.plan(frames, {"wg_per_cu": 1, "frames_per_job": 1, "order": 1, "split_edges": split})
        check(ctxs, frames, who, ref, run_stream(sc, frames, llr, hb))


def test_stream_pairs_of_short_frames():
    """Next to normal frames the 12-wave kernel decodes two short frames of a
    code per workgroup, one per lane: every frame must still equal its own
    context's decode (learned with and without fallback, frames that fail,
    an odd frame count per code) and the unpaired stream's."""
    from cusat_amd import learn
    torch = _torch()
    mix = [(0, "3/4", "8PSK", 7.9), (1, "1/2", "QPSK", 1.9), (1, "3/5", "8PSK", 6.2),
           (1, "5/6", "16APSK", 11.4), (1, "1/4", "QPSK", -0.6), (1, "2/3", "QPSK", 3.3)]
    # pairs are planned only when the frames outnumber the workgroups (one per CU)
    per_code = (torch.cuda.get_device_properties(0).multi_processor_count // len(mix) + 5) | 1
    ctxs = [ca.FecContext(fs, r, m, max_trials=50, max_batch=per_code) for fs, r, m, _ in mix]
    path = learn.params_for(1, "1/2", "QPSK")
    _, bq, aq = learn.load_params(path)
    ctxs[1].set_params(bq, aq)
    ctxs[1]._learned = True
    ctxs, frames, llr, by_code, who, hb = build_stream(mix, per_code, 11, contexts=ctxs)
    ref = per_code_reference(ctxs, by_code, learned_trials=bq.shape[0])
    t = np.concatenate([r[1] for r in ref[1:]])
    assert (t < 0).any() and (t >= 0).any(), "the short codes need failing and converging frames"
    assert len(set(ref[1][1].tolist())) > 1, "the learned code's frames need different trial counts"
    for flags in (ca.STREAM_LEARNED, ca.STREAM_LEARNED | ca.STREAM_NO_PAIRS):
        sc = ca.StreamContext(ctxs, len(frames), flags)
        for tuning in ({"wg_per_cu": 1, "frames_per_job": 1, "order": 1, "split_edges": 1},
                       {"wg_per_cu": 1, "frames_per_job": 3, "order": 1, "split_edges": 1},
                       {"wg_per_cu": 1, "frames_per_job": 2, "order": 0, "split_edges": 1}):
            sc.plan(frames, tuning)
            check(ctxs, frames, who, ref, run_stream(sc, frames, llr, hb))
        sc.close()


def test_stream_graph_replays():
    torch = _torch()
    mix = MIX[:3] + MIX[5:6]
    ctxs, frames, llr, by_code, who, hb = build_stream(mix, 4, 11)
    sc = ca.StreamContext(ctxs, len(frames))
    sc.plan(frames)
    n = len(frames)
    hard = torch.zeros(hb, dtype=torch.uint8, device="cuda")
    tr = torch.zeros(n, dtype=torch.int32, device="cuda")
    va = torch.zeros(n, dtype=torch.int32, device="cuda")
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        sc.decode(llr, hard, tr, va)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        sc.decode(llr, hard, tr, va)
    for rep in range(3):
        # new LLRs in the same buffer, written eagerly between replays
        _, _, _, by_new, _, _ = build_stream(mix, 4, 20 + rep, contexts=ctxs)
        frames_new = frames  # same layout: the new frames go to the old slots
        parts = []
        for (k, j) in who:
            parts.append(by_new[k][j])
        llr.copy_(torch.cat(parts))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = (hard.cpu().numpy(), tr.cpu().numpy(), va.cpu().numpy(), np.zeros(n, np.int32))
        ref = per_code_reference(ctxs, by_new)
        check(ctxs, frames_new, who, ref, got)


def _subset(frames, who, pick, ctxs):
    """Frame records of the frames `pick` (indices into frames), packed into
    fresh llr / hard positions in the given order, and the slices of the
    original llr buffer to gather."""
    sub = np.zeros(len(pick), ca.STREAM_FRAME)
    lo = ho = 0
    src = []
    for n, p in enumerate(pick):
        k = who[p][0]
        sub[n] = (k, 0, lo, ho)
        src.append((int(frames[p]["llr_offset"]), ctxs[k].info["n_ldpc"]))
        lo += ctxs[k].info["n_ldpc"]
        ho += ctxs[k].info["k_ldpc"] // 8
    return sub, src, lo, ho


def test_stream_graph_replays_new_plans():
    """A captured launch replays whatever plan cusat_stream_plan_async made
    current: fewer frames, other codes, another order, an empty plan."""
    torch = _torch()
    mix = MIX[:3] + MIX[5:7]
    ctxs, frames, llr, by_code, who, hb = build_stream(mix, 4, 12)
    ref = per_code_reference(ctxs, by_code)
    n = len(frames)
    sc = ca.StreamContext(ctxs, n)
    sc.plan(frames)
    llr_w = torch.zeros_like(llr)
    hard = torch.zeros(hb, dtype=torch.uint8, device="cuda")
    tr = torch.zeros(n, dtype=torch.int32, device="cuda")
    va = torch.zeros(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        sc.decode(llr_w, hard, tr, va)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        sc.decode(llr_w, hard, tr, va)
    rng = np.random.default_rng(5)
    picks = [rng.permutation(n)[:7], np.arange(n)[::-1], np.array([], np.int64), rng.permutation(n)]
    for pick in picks:
        sub, src, lo, ho = _subset(frames, who, pick, ctxs)
        if len(pick):
            llr_w[:lo].copy_(torch.cat([llr[o:o + ln] for o, ln in src]))
        tr.fill_(-7)
        va.fill_(-7)
        hard.zero_()
        sc.plan_async(sub)  # on the current stream, before the replay
        g.replay()
        torch.cuda.synchronize()
        t, v, h = tr.cpu().numpy(), va.cpu().numpy(), hard.cpu().numpy()
        for m, p in enumerate(pick):
            k, j = who[p]
            kb = ctxs[k].info["k_ldpc"] // 8
            assert t[m] == ref[k][1][j] and v[m] == ref[k][2][j], (m, k, j)
            np.testing.assert_array_equal(h[sub[m]["hard_offset"]:sub[m]["hard_offset"] + kb], ref[k][0][j])
        assert (t[len(pick):] == -7).all()  # frames outside the plan untouched


def test_stream_plan_async_back_to_back():
    """More plans in flight on one stream than the pinned staging ring
    holds, with no host synchronisation between plan and decode: every
    decode runs the plan made just before it."""
    torch = _torch()
    mix = MIX[:2] + MIX[5:6]
    ctxs, frames, llr, by_code, who, hb = build_stream(mix, 3, 13)
    ref = per_code_reference(ctxs, by_code)
    n = len(frames)
    sc = ca.StreamContext(ctxs, n)
    rng = np.random.default_rng(9)
    outs = []
    for r in range(9):
        pick = rng.permutation(n)[: 1 + r % n]
        sub, src, lo, ho = _subset(frames, who, pick, ctxs)
        lw = torch.cat([llr[o:o + ln] for o, ln in src])
        h = torch.zeros(ho, dtype=torch.uint8, device="cuda")
        t = torch.full((len(pick),), -7, dtype=torch.int32, device="cuda")
        v = torch.full((len(pick),), -7, dtype=torch.int32, device="cuda")
        sc.plan_async(sub)
        sc.decode(lw, h, t, v)
        outs.append((pick, sub, h, t, v, lw))
    torch.cuda.synchronize()
    for pick, sub, h, t, v, _ in outs:
        t, v, h = t.cpu().numpy(), v.cpu().numpy(), h.cpu().numpy()
        for m, p in enumerate(pick):
            k, j = who[p]
            kb = ctxs[k].info["k_ldpc"] // 8
            assert t[m] == ref[k][1][j] and v[m] == ref[k][2][j], (m, k, j)
            np.testing.assert_array_equal(h[sub[m]["hard_offset"]:sub[m]["hard_offset"] + kb], ref[k][0][j])
