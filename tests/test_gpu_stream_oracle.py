"""The persistent stream decoder against the oracle on the mixed ACM
workloads, through the receive path (cusat_stream_rx_plan_async, then the
three launches of cusat_stream_rx one by one).

* mixed12: the 12 DVB-S2 MODCODs of the 12-MODCOD stream (QPSK 1/4, 1/2,
  2/3, 3/4; 8PSK 3/5, 3/4, 9/10; 16APSK 2/3, 3/4, 5/6; 32APSK 3/4, 9/10),
  normal and short FECFRAMEs (short where the code exists), Es/N0 at the
  MODCOD's threshold (lib/cusat_dsp/acm_config.h:54-121) + 0.5 dB.
* s2x: the same plus 64APSK 4/5 (8+16+20+20 rings), 128APSK 3/4 and
  256APSK 3/4 (normal only), threshold + 0.3 dB.

Small mixes (4 frames of every (frame size, MODCOD), arrival order
shuffled): every frame's LLRs equal the oracle's demap, and its hard bits,
trials, BCH flag, fallback flag and data-field status equal the oracle's
decode + BCH check -- plain mode, and learned + fallback mode with the
trained parameters of cusat_amd/data/learned composed from oracle calls
(learned decode with the per-(iteration, layer) offsets and scales, BCH,
then the plain decoder on every frame failing frame_passes_fec_checks,
rx_fec_chain.h:193-196).  Full-size streams (2048 mixed12 frames, 4096 s2x
frames, the composition of stream_bench.make_stream): every frame against
the oracle, the batched per-MODCOD launches and the result properties.
"""
import numpy as np
import pytest

import cusat_amd as ca
from cusat_amd import learn, stream_bench
from oracle import oracle_ctypes as oc
from tests import bb_ref
from tests.fec_util import oracle_code, oracle_demap, snr_scales

pytestmark = pytest.mark.gpu

T = 50


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def workload_groups(workload):
    mods, margin = (learn.MODCODS12, 0.5) if workload == "mixed12" else (learn.MODCODS12 + learn.S2X, 0.3)
    return [(fs, rate, mod, th + margin) for rate, mod, th in mods
            for fs in ((0, 1) if learn.has_short(rate, mod) else (0,))]


class Stream:
    """A mixed stream: contexts per (frame size, MODCOD, Es/N0) group,
    host symbols per frame, the rx records in arrival order."""

    def __init__(self, spec, seed, learned):
        torch = _torch()
        self.keys = sorted(set(spec))
        idx = {k: i for i, k in enumerate(self.keys)}
        count = {k: 0 for k in self.keys}
        for s in spec:
            count[s] += 1
        self.ctxs, self.sym, self.params, self._prbs = [], {}, {}, {}
        for gi, k in enumerate(self.keys):
            fs, rate, mod, es = k
            c = ca.FecContext(fs, rate, mod, max_trials=T, max_batch=1)
            self.ctxs.append(c)
            self.sym[k], _, _ = ca.gen_frames(fs, rate, mod, seed + 1000 * gi, count[k], es)
            path = learn.params_for(fs, rate, mod) if learned else None
            if path:
                _, bq, aq = learn.load_params(path)
                self.params[gi] = (bq, aq)
        # arrival order: the spec's; frame n is row j of its group
        slot = {k: 0 for k in self.keys}
        self.who = []
        for s in spec:
            self.who.append((idx[s], slot[s]))
            slot[s] += 1
        n = len(spec)
        self.recs = np.zeros(n, ca.STREAM_RX_FRAME)
        parts, so, oo = [], 0, 0
        self.llr_off, self.hard_off = np.zeros(n, np.int64), np.zeros(n, np.int64)
        lo = ho = 0
        for f, (gi, j) in enumerate(self.who):
            k = self.keys[gi]
            inf = self.ctxs[gi].info
            nv, g = snr_scales(k[3])
            self.recs[f] = (gi, 0, so, oo, nv, g)
            parts.append(self.sym[k][j])
            so += self.sym[k].shape[1]
            oo += inf["k_bch"] // 8 - 10
            self.llr_off[f], self.hard_off[f] = lo, ho
            lo += inf["n_ldpc"]
            ho += inf["k_ldpc"] // 8
        self.n, self.out_bytes, self.llr_total, self.hard_total = n, oo, lo, ho
        self.d_sym = torch.from_numpy(np.ascontiguousarray(np.concatenate(parts))).cuda()
        for gi, (bq, aq) in self.params.items():
            self.ctxs[gi].set_params(bq, aq)

    def run(self, learned, tuning=None):
        """rx plan + the three launches; returns trials, valid, fallback,
        dfl and the host copies of the LLR and hard work buffers."""
        torch = _torch()
        sc = ca.StreamContext(self.ctxs, self.n, ca.STREAM_LEARNED if learned else 0)
        if tuning is not None:
            sc.plan(np.zeros(0, ca.STREAM_FRAME), tuning)
        out = torch.zeros(self.out_bytes + 1, dtype=torch.uint8, device="cuda")
        tr, va, fb, dfl = (torch.full((self.n,), -7, dtype=torch.int32, device="cuda") for _ in range(4))
        sc.rx_plan_async(self.recs)
        sc.rx_demap(self.d_sym)
        sc.rx_decode(tr, va, fb)
        sc.rx_bb(out, tr, va, dfl)
        torch.cuda.synchronize()
        a_llr, a_hard = sc.rx_work()
        llr = ca.copy_from_device(a_llr, 2 * self.llr_total).view(np.int16)
        hard = ca.copy_from_device(a_hard, self.hard_total)
        res = [x.cpu().numpy() for x in (tr, va, fb, dfl)]
        sc.close()
        return res + [llr, hard]

    def group_llr(self, llr, gi):
        """LLRs of group gi's frames, rows in group order, from the work buffer."""
        N = self.ctxs[gi].info["n_ldpc"]
        rows = [(j, f) for f, (g, j) in enumerate(self.who) if g == gi]
        out = np.zeros((len(rows), N), np.int16)
        for j, f in rows:
            out[j] = llr[self.llr_off[f]:self.llr_off[f] + N]
        return out

    def oracle(self, llr, learned, nthreads=16):
        """Per frame (arrival order): hard bytes, trials, BCH flag, fallback
        flag from the oracle, on the LLRs the GPU demapped."""
        hard, tr, va, fb = [None] * self.n, np.zeros(self.n, np.int32), np.zeros(self.n, np.int32), \
            np.zeros(self.n, np.int32)
        for gi, k in enumerate(self.keys):
            fs, rate = k[0], k[1]
            inf = self.ctxs[gi].info
            L = self.group_llr(llr, gi)
            code = oracle_code(fs, rate)
            h, t = code.decode(L, T, nthreads=nthreads)
            v = oc.bch_check(h, inf["n_bch"], fs, inf["bch_t"])
            redo = np.zeros(len(L), bool)
            if learned and gi in self.params:
                bq, aq = self.params[gi]
                tl = bq.shape[0]
                h_l, t_l = code.decode(L, tl, beta=bq, alpha=aq, n_param_iters=tl, nthreads=nthreads)
                v_l = oc.bch_check(h_l, inf["n_bch"], fs, inf["bch_t"])
                redo = ~((t_l >= 0) & (v_l == 1))
                h = np.where(redo[:, None], h, h_l)
                t = np.where(redo, t, t_l)
                v = np.where(redo, v, v_l)
            for f, (g, j) in enumerate(self.who):
                if g == gi:
                    hard[f], tr[f], va[f], fb[f] = h[j], t[j], v[j], int(redo[j])
        return hard, tr, va, fb

    def check(self, got, want, learned):
        tr, va, fb, dfl, _, hard = got
        h_o, t_o, v_o, f_o = want
        for f, (gi, j) in enumerate(self.who):
            kb = self.ctxs[gi].info["k_ldpc"] // 8
            assert tr[f] == t_o[f], (f, self.keys[gi], j)
            assert va[f] == v_o[f], (f, self.keys[gi], j)
            assert fb[f] == (f_o[f] if learned else 0), (f, self.keys[gi], j)
            np.testing.assert_array_equal(hard[self.hard_off[f]:self.hard_off[f] + kb], h_o[f])
            # data-field status: -1 for frames failing the FEC checks, else
            # what the BBHEADER of the descrambled frame says (random info
            # bits: almost always no valid header, 0)
            if t_o[f] >= 0 and v_o[f] == 1:
                kbch = self.ctxs[gi].info["k_bch"] // 8
                if gi not in self._prbs:
                    self._prbs[gi] = ca.bb_prbs(kbch)
                ext = bb_ref.extract(h_o[f][:kbch] ^ self._prbs[gi], kbch)
                assert dfl[f] == (len(ext) if ext else 0), f
            else:
                assert dfl[f] == -1, f


def small_spec(workload, per_group, seed):
    groups = workload_groups(workload)
    spec = [g for g in groups for _ in range(per_group)]
    rng = np.random.default_rng(seed)
    return [spec[p] for p in rng.permutation(len(spec))]


@pytest.mark.parametrize("workload,learned", [("mixed12", False), ("s2x", False), ("s2x", True),
                                              ("mixed12", True)])
def test_stream_small_mix_vs_oracle(workload, learned):
    spec = small_spec(workload, 4, 11 if workload == "mixed12" else 13)
    st = Stream(spec, 31 if workload == "mixed12" else 37, learned)
    if learned:
        assert len(st.params) >= 8, "trained parameters for most codes of the workload"
    got = st.run(learned)
    # LLRs: the mixed demap equals the oracle's demap of every frame
    for f, (gi, j) in enumerate(st.who):
        k = st.keys[gi]
        nv, g = snr_scales(k[3])
        want = oracle_demap(st.ctxs[gi], st.sym[k][j:j + 1], nv, g)[0]
        N = st.ctxs[gi].info["n_ldpc"]
        np.testing.assert_array_equal(got[4][st.llr_off[f]:st.llr_off[f] + N], want, err_msg=str(k))
    want = st.oracle(got[4], learned)
    st.check(got, want, learned)
    tr = got[0]
    assert (tr >= 0).sum() >= len(tr) // 2, "most frames converge above threshold"
    if learned:
        assert got[2].sum() < len(tr), "not every frame falls back"


def test_stream_small_mix_tunings_vs_oracle():
    """The 6-wave build and other job shapes give the same frames."""
    spec = small_spec("s2x", 4, 17)
    st = Stream(spec, 41, True)
    want = None
    for tu in ({"wg_per_cu": 1, "frames_per_job": 2, "order": 0, "split_edges": 0},
               {"wg_per_cu": 1, "frames_per_job": 4, "order": 1, "split_edges": 1}):
        got = st.run(True, tu)
        if want is None:
            want = st.oracle(got[4], True)
        st.check(got, want, True)


@pytest.mark.parametrize("workload,n_frames,learned", [("mixed12", 2048, False), ("s2x", 4096, True)])
def test_stream_full_size_vs_oracle(workload, n_frames, learned):
    torch = _torch()
    spec = stream_bench.make_stream(workload, n_frames, 20240612)
    st = Stream(spec, 20240612, learned)
    got = st.run(learned)
    tr, va, fb, dfl, llr, hard = got
    # properties of the result
    assert ((tr == -1) | ((tr >= 0) & (tr <= T))).all()
    assert set(np.unique(va)) <= {0, 1} and set(np.unique(fb)) <= {0, 1}
    acc = (tr >= 0) & (va == 1)
    assert acc.mean() > 0.99, "threshold + margin: nearly every frame decodes"
    if not learned:
        assert not fb.any()
    # every frame against the oracle
    st.check(got, st.oracle(llr, learned), learned)
    # and against the batched per-MODCOD launches on the same LLRs
    for gi, k in enumerate(st.keys):
        c = st.ctxs[gi]
        L = st.group_llr(llr, gi)
        B = len(L)
        big = ca.FecContext(k[0], k[1], k[2], max_trials=T, max_batch=B)
        d = torch.from_numpy(L).cuda()
        h = torch.zeros((B, c.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
        t, v, f = (torch.zeros(B, dtype=torch.int32, device="cuda") for _ in range(3))
        if gi in st.params:
            big.set_params(*st.params[gi])
            big.rx_fec_learned(d, h, t, v, f, B, st.params[gi][0].shape[0])
        else:
            big.decode(d, h, t, B)
            big.bch_check(h, v, B)
        torch.cuda.synchronize()
        t, v = t.cpu().numpy(), v.cpu().numpy()
        for f_, (g, j) in enumerate(st.who):
            if g == gi:
                assert (tr[f_], va[f_]) == (t[j], v[j]), (k, j)


def test_set_params_after_async_plan():
    """A stream planned (async) before its contexts' learned parameters
    change decodes with the new ones from the next plan call on: a larger
    table (its buffer reallocated), then a cleared one."""
    torch = _torch()
    fs, rate, mod, es, B = 0, "2/3", "16APSK", 9.1, 24
    ctx = ca.FecContext(fs, rate, mod, max_trials=T, max_batch=B)
    sym, _, _ = ca.gen_frames(fs, rate, mod, 5, B, es)
    nv, g = snr_scales(es)
    llr = torch.empty((B, ctx.info["n_ldpc"]), dtype=torch.int16, device="cuda")
    ctx.demap(torch.from_numpy(sym).cuda(), llr, B, nv, g)
    q = ctx.info["q"]
    rng = np.random.default_rng(9)
    sc = ca.StreamContext([ctx], B, ca.STREAM_LEARNED)
    frames = np.zeros(B, ca.STREAM_FRAME)
    for b in range(B):
        frames[b] = (0, 0, b * ctx.info["n_ldpc"], b * (ctx.info["k_ldpc"] // 8))
    ctx.set_params(np.ones((4, q), np.int16), rng.integers(58, 70, (4, q)).astype(np.int16))
    sc.plan_async(frames)
    code = oracle_code(fs, rate)
    L = llr.cpu().numpy()
    for n_it in (12, 0):
        bq = rng.integers(0, 3, (n_it, q)).astype(np.int16)
        aq = rng.integers(56, 72, (n_it, q)).astype(np.int16)
        ctx.set_params(bq, aq)
        sc.plan_async(frames)
        hard = torch.zeros(B * (ctx.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
        tr, va, fb = (torch.full((B,), -7, dtype=torch.int32, device="cuda") for _ in range(3))
        sc.decode(llr, hard, tr, va, fb)
        torch.cuda.synchronize()
        if n_it:
            h_l, t_l = code.decode(L, n_it, beta=bq, alpha=aq, n_param_iters=n_it)
            v_l = oc.bch_check(h_l, ctx.info["n_bch"], fs, ctx.info["bch_t"])
            redo = ~((t_l >= 0) & (v_l == 1))
            h_p, t_p = code.decode(L, T)
            want_t = np.where(redo, t_p, t_l)
            want_h = np.where(redo[:, None], h_p, h_l)
        else:
            # cleared: every code decodes plain, nothing falls back
            want_h, want_t = code.decode(L, T)
            redo = np.zeros(B, bool)
        np.testing.assert_array_equal(tr.cpu().numpy(), want_t)
        np.testing.assert_array_equal(fb.cpu().numpy(), redo.astype(np.int32))
        np.testing.assert_array_equal(hard.cpu().numpy().reshape(B, -1), want_h)
