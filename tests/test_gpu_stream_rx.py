"""Mixed-MODCOD receive batch (cusat_stream_rx_plan_async + cusat_stream_rx)
on the GPU.

A micro-batch of XFECFRAMEs of several MODCODs (normal and short frames;
QPSK / 8PSK formulas and APSK tables, 8APSK included), each with its own
demapper scaling, symbols and data-field slots at arbitrary offsets, goes
through demap -> decode -> BCH -> BB in three launches.  Per frame, trials,
BCH flag, data-field length and bytes must equal what the frame's own
context gives (cusat_rx_fec_batch + cusat_bb_extract, one frame at a time,
same scaling), whatever the tuning; frames above threshold return their GSE
payloads; a captured launch replays new batches.
"""
import numpy as np
import pytest

import cusat_amd as ca
from tests import bb_ref
from tests.fec_util import dd_snr_host as _dd_host, snr_scales

pytestmark = pytest.mark.gpu

# (framesize, rate, modulation, Es/N0): near threshold, well above, one far
# below (every frame fails)
MIX = [(0, "1/2", "QPSK", 1.6), (0, "3/4", "8PSK", 8.6), (0, "2/3", "16APSK", 9.4),
       (0, "3/4", "32APSK", 13.3), (0, "100/180", "8APSK", 5.6), (1, "1/2", "QPSK", 2.6),
       (1, "5/6", "16APSK", 11.8), (0, "9/10", "8PSK", 9.0)]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make_batch(mix, per_code, seed):
    """Contexts, per-frame (code, symbols, noise_var, llr_gain, payload),
    shuffled into arrival order."""
    rng = np.random.default_rng(seed)
    ctxs = [ca.FecContext(fs, r, m, max_trials=50, max_batch=1) for fs, r, m, _ in mix]
    frames = []
    for k, (fs, r, m, es) in enumerate(mix):
        kbch = ctxs[k].info["k_bch"] // 8
        prbs = ca.bb_prbs(kbch)
        lens = rng.integers(0, kbch - 10, per_code)
        lens[0] = kbch - 10
        if per_code > 1:
            lens[1] = 0
        pay = [rng.integers(0, 256, n, dtype=np.uint8) for n in lens]
        info = np.stack([bb_ref.make_bbframe(kbch, p, rng=rng) ^ prbs for p in pay])
        sym, _ = ca.encode_frames(fs, r, m, seed * 50 + k, info, es)
        for j in range(per_code):
            nv, g = snr_scales(es + rng.uniform(-0.3, 0.3))  # every frame its own scaling
            frames.append((k, sym[j], nv, g, pay[j]))
    return ctxs, [frames[p] for p in rng.permutation(len(frames))]


def reference(ctxs, frames):
    """Per frame: its own context, one frame at a time."""
    torch = _torch()
    out = []
    for k, sym, nv, g, _ in frames:
        c = ctxs[k]
        info = c.info
        llr = torch.empty((1, info["n_ldpc"]), dtype=torch.int16, device="cuda")
        hard = torch.empty((1, info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
        tr = torch.empty(1, dtype=torch.int32, device="cuda")
        va = torch.empty(1, dtype=torch.int32, device="cuda")
        o = torch.zeros((1, info["k_bch"] // 8 - 10), dtype=torch.uint8, device="cuda")
        dfl = torch.empty(1, dtype=torch.int32, device="cuda")
        c.rx_fec_batch(torch.from_numpy(np.ascontiguousarray(sym[None])).cuda(), llr, 1, nv, g, hard, tr, va)
        c.bb_extract(hard, tr, va, o, dfl, 1)
        torch.cuda.synchronize()
        n = int(dfl.item())
        out.append((int(tr.item()), int(va.item()), n, o.cpu().numpy()[0, :max(n, 0)].tobytes()))
    return out


def pack(ctxs, frames, rng):
    """Device symbol buffer and STREAM_RX_FRAME records with gaps between
    frames; out-slot offsets likewise."""
    torch = _torch()
    recs = np.zeros(len(frames), ca.STREAM_RX_FRAME)
    so = oo = 0
    parts = []
    for n, (k, sym, nv, g, _) in enumerate(frames):
        gap = int(rng.integers(0, 40))
        parts.append(np.zeros((gap, 2), np.float32))
        so += gap
        oo += int(rng.integers(0, 100))
        recs[n] = (k, 0, so, oo, nv, g)
        parts.append(sym)
        so += sym.shape[0]
        oo += ctxs[k].info["k_bch"] // 8 - 10
    d_sym = torch.from_numpy(np.ascontiguousarray(np.concatenate(parts))).cuda()
    return recs, d_sym, oo


def outputs(n, out_bytes):
    torch = _torch()
    out = torch.full((out_bytes + 1,), 0xEE, dtype=torch.uint8, device="cuda")
    tr = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    va = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    dfl = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    return out, tr, va, dfl


def check(frames, recs, ref, out, tr, va, dfl, above_threshold=True):
    out, tr, va, dfl = (x.cpu().numpy() for x in (out, tr, va, dfl))
    for n, (k, _, _, _, pay) in enumerate(frames):
        t0, v0, d0, b0 = ref[n]
        assert (tr[n], va[n], dfl[n]) == (t0, v0, d0), (n, k)
        o = recs[n]["out_offset"]
        assert out[o:o + max(d0, 0)].tobytes() == b0, (n, k)
        if d0 > 0:
            assert b0 == pay.tobytes(), (n, k)
    return dfl


@pytest.mark.parametrize("tuning", [None, {"wg_per_cu": 1, "frames_per_job": 2, "order": 0, "split_edges": 0},
                                    {"wg_per_cu": 1, "frames_per_job": 1, "order": 1, "split_edges": 1}])
def test_stream_rx_matches_per_frame_chain(tuning):
    torch = _torch()
    ctxs, frames = make_batch(MIX, 4, 21)
    ref = reference(ctxs, frames)
    sc = ca.StreamContext(ctxs, len(frames))
    if tuning is not None:
        sc.plan(np.zeros(0, ca.STREAM_FRAME), tuning)
    recs, d_sym, ob = pack(ctxs, frames, np.random.default_rng(2))
    out, tr, va, dfl = outputs(len(frames), ob)
    sc.rx_plan_async(recs)
    sc.rx(d_sym, out, tr, va, dfl)
    torch.cuda.synchronize()
    d = check(frames, recs, ref, out, tr, va, dfl)
    # the mix covers all three outcomes
    assert (d == -1).any() and (d == 0).any() and (d > 0).sum() >= len(frames) // 2
    # bytes outside the frames' data fields untouched
    o = out.cpu().numpy()
    mask = np.ones(ob + 1, bool)
    for n in range(len(frames)):
        mask[recs[n]["out_offset"]:recs[n]["out_offset"] + max(d[n], 0)] = False
    assert (o[mask] == 0xEE).all()


def test_stream_rx_graph_replays_new_batches():
    torch = _torch()
    ctxs, frames = make_batch(MIX[:3] + MIX[5:7], 3, 22)
    ref = reference(ctxs, frames)
    n = len(frames)
    sc = ca.StreamContext(ctxs, n)
    rng = np.random.default_rng(4)
    recs, d_sym, ob = pack(ctxs, frames, rng)
    sym_w = torch.zeros(d_sym.shape[0] + 4096, 2, dtype=torch.float32, device="cuda")
    out, tr, va, dfl = outputs(n, ob + 4096)
    sc.rx_plan_async(np.zeros(0, ca.STREAM_RX_FRAME))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        sc.rx(sym_w, out, tr, va, dfl)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        sc.rx(sym_w, out, tr, va, dfl)
    for rep in range(3):
        pick = rng.permutation(n)[: n - 2 * rep]
        sub = [frames[p] for p in pick]
        r2, d2, ob2 = pack(ctxs, sub, rng)
        sym_w[: d2.shape[0]].copy_(d2)
        out.fill_(0xEE)
        for x in (tr, va, dfl):
            x.fill_(-7)
        sc.rx_plan_async(r2)
        g.replay()
        torch.cuda.synchronize()
        check(sub, r2, [ref[p] for p in pick], out, tr, va, dfl)
        assert (tr[len(pick):].cpu().numpy() == -7).all()


def test_stream_rx_rejects_bad_records():
    _torch()
    ctxs, frames = make_batch(MIX[:1], 1, 23)
    sc = ca.StreamContext(ctxs, 2)
    recs = np.zeros(1, ca.STREAM_RX_FRAME)
    recs[0] = (1, 0, 0, 0, 1.0, 4.0)  # no such code
    with pytest.raises(RuntimeError):
        sc.rx_plan_async(recs)
    recs[0] = (0, 0, 0, 0, 0.0, 4.0)  # noise variance must be positive
    with pytest.raises(RuntimeError):
        sc.rx_plan_async(recs)
    with pytest.raises(RuntimeError):
        sc.rx_plan_async(np.zeros(3, ca.STREAM_RX_FRAME))  # more than max_frames


def test_stream_rx_scaling_from_the_device_snr_state():
    """Frames flagged RX_SCALE_STATE demap with their code's SNR state after
    the batch's events: a host value, or the GPU estimate (checked against a
    float32 restatement of the host estimate), applied in order."""
    torch = _torch()
    mix = [(0, "3/4", "8PSK", 8.6), (1, "1/2", "QPSK", 2.6), (1, "5/6", "16APSK", 11.8)]
    ctxs, frames = make_batch(mix, 3, 31)
    n = len(frames)
    sc = ca.StreamContext(ctxs, n)
    recs, d_sym, ob = pack(ctxs, frames, np.random.default_rng(8))
    recs["flags"] = ca.RX_SCALE_STATE
    recs["noise_var"] = 0.0  # ignored for flagged frames
    # code 0: host value; code 1: estimate from its first frame; code 2: an
    # estimate followed by a host value (the host value wins)
    first = {k: next(i for i, f in enumerate(frames) if f[0] == k) for k in range(3)}
    v0 = np.float32(10.0 ** 0.86)
    v2 = np.float32(10.0 ** 1.18)
    ev = np.zeros(4, ca.STREAM_RX_EVENT)
    ev[0] = (0, -1, v0, 8.6)
    ev[1] = (1, first[1], 0.0, 0.0)
    ev[2] = (2, first[2], 0.0, 0.0)
    ev[3] = (2, -1, v2, 11.8)
    state = torch.tensor([[1.0, 4.0, 0.0, 0.0]] * 3, dtype=torch.float32, device="cuda")
    state_out = torch.zeros_like(state)
    out, tr, va, dfl = outputs(n, ob)
    sc.rx_plan_async(recs, ev)
    sc.rx_dd(d_sym, state, state_out)
    sc.rx(d_sym, out, tr, va, dfl)
    torch.cuda.synchronize()
    st = state_out.cpu().numpy()
    est1 = _dd_host(frames[first[1]][1], None, True)  # QPSK: sign decisions
    want_lin = [v0, est1 if est1 > 1 else np.float32(0.0), v2]
    for k in range(3):
        assert st[k, 2].view(np.uint32) == np.float32(want_lin[k]).view(np.uint32), (k, st[k], want_lin[k])
        nv = np.float32(np.float32(1.0) / np.float32(want_lin[k]))
        assert st[k, 0].view(np.uint32) == nv.view(np.uint32)
        assert st[k, 1] == min(np.float32(np.float32(4.0) / nv), np.float32(1000.0))
    assert np.isnan(st[1, 3]) and st[0, 3] == np.float32(8.6) and st[2, 3] == np.float32(11.8)
    # the frames were demapped with these states: per-frame references at
    # the same (noise_var, llr_gain) give the same outputs
    ref = reference(ctxs, [(k, sym, float(st[k, 0]), float(st[k, 1]), pay) for k, sym, _, _, pay in frames])
    check(frames, recs, ref, out, tr, va, dfl)
    # the device state itself was updated for the next batch
    assert np.array_equal(state.cpu().numpy(), st, equal_nan=True)
    # the estimate and resolve launches on their own give the same state
    state2 = torch.tensor([[1.0, 4.0, 0.0, 0.0]] * 3, dtype=torch.float32, device="cuda")
    state2_out = torch.zeros_like(state2)
    sc.rx_dd_estimate(d_sym)
    sc.rx_dd_resolve(state2, state2_out)
    torch.cuda.synchronize()
    assert np.array_equal(state2_out.cpu().numpy(), st, equal_nan=True)
