"""Mixed ACM stream benchmark: one persistent launch (hipGraph replay) vs one
batched launch per MODCOD, on the same packed int16 LLRs.

Workloads (synthetic, seeded; AWGN at each MODCOD's threshold + margin, the
thresholds of lib/cusat_dsp/acm_config.h:54-121, normal-frame values also
used for short frames):

* mixed12: 2048 frames drawn uniformly from 12 DVB-S2 MODCODs (QPSK 1/4,
  1/2, 3/4; 8PSK 3/5, 9/10; 16APSK 2/3, 5/6; 32APSK 3/4, 9/10; plus QPSK 2/3,
  8PSK 3/4, 16APSK 3/4 to make twelve), 30% short FECFRAMEs where the MODCOD
  has one (no 9/10 short code), threshold + 0.5 dB, plain decoder.
* s2x: 4096 frames over 64APSK 4/5, 128APSK 3/4, 256APSK 3/4 and the twelve
  above, 25% short (DVB-S2 MODCODs only), threshold + 0.3 dB; every
  (frame size, MODCOD) with trained parameters (cusat_amd/data/learned, one
  file per code from `python -m cusat_amd.learn train-set`) runs the learned
  decoder with in-kernel plain fallback, against the all-plain stream on the
  same LLRs.

    python -m cusat_amd.stream_bench --workload mixed12
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from cusat_amd.learn import MODCODS12, NO_SHORT, NO_SHORT_MOD, S2X


def snr_scales(es):
    s = np.float32(10.0) ** (np.float32(es) / np.float32(10.0))
    nv = np.float32(1.0) / s
    return float(nv), float(min(np.float32(4.0) / nv, np.float32(1000.0)))


def make_stream(workload, n_frames, seed, log=print):
    """Frames of the workload: list of (framesize, rate, modulation, Es/N0)
    per frame, in stream order."""
    rng = np.random.default_rng(seed)
    if workload == "mixed12":
        mods, short_p, margin = MODCODS12, 0.30, 0.5
    else:
        mods, short_p, margin = MODCODS12 + S2X, 0.25, 0.3
    # short-frame probability among the MODCODs that have a short code, so
    # that short_p of all frames are short
    eligible = [m for m in mods if m[0] not in NO_SHORT and m[1] not in NO_SHORT_MOD]
    p_el = min(1.0, short_p * len(mods) / len(eligible))
    out = []
    for _ in range(n_frames):
        rate, mod, th = mods[rng.integers(len(mods))]
        can_short = rate not in NO_SHORT and mod not in NO_SHORT_MOD
        fs = 1 if (can_short and rng.random() < p_el) else 0
        out.append((fs, rate, mod, th + margin))
    return out


_POPCOUNT = np.array([bin(v).count("1") for v in range(256)], np.uint8)


def bbheader(kbch_bytes):
    """BBHEADER (EN 302 307-1 clause 5.1.6, table 3) of a frame whose data
    field fills the BBFRAME: MATYPE-1 = generic packetized TS/GS, single
    input stream, CCM, roll-off 0.20; UPL 0, DFL = 8 (k_bch/8 - 10) bits,
    SYNC 0, SYNCD 0xFFFF, then the CRC-8 of the first nine bytes."""
    h = np.zeros(10, np.uint8)
    dfl = 8 * (kbch_bytes - 10)
    h[0] = (1 << 5) | (1 << 4) | 2
    h[4], h[5] = dfl >> 8, dfl & 0xFF
    h[7] = h[8] = 0xFF
    import cusat_amd as ca
    h[9] = ca.bb_crc8(h[:9])
    return h


class RxStream:
    """A mixed ACM stream for the receive path (cusat_stream_rx): frames of
    make_stream's composition, each a scrambled BBFRAME (bbheader + a
    data field of seeded random bytes) through the GPU transmit chain
    (cusat_gen_frames_dev: BCH, LDPC, interleave + map, AWGN at the frame's
    Es/N0), symbols packed back to back in arrival order in HBM.  Holds the
    sent data fields and codewords to score what the receiver delivers."""

    def __init__(self, workload="s2x", n_frames=None, seed=20240612, learned=True, max_trials=50):
        import torch
        import cusat_amd as ca
        from cusat_amd import learn
        n_frames = n_frames or (2048 if workload == "mixed12" else 4096)
        spec = make_stream(workload, n_frames, seed)
        self.spec = spec
        self.keys = sorted(set(spec))
        idx = {k: i for i, k in enumerate(self.keys)}
        count = {k: 0 for k in self.keys}
        for s_ in spec:
            count[s_] += 1
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        self.ctxs, self.learned, self.params = [], set(), {}
        sym_g, pay_g, cw_g = [], [], []
        for gi, k in enumerate(self.keys):
            fs, rate, mod, es = k
            c = ca.FecContext(fs, rate, mod, max_trials=max_trials, max_batch=count[k])
            inf = c.info
            kb = inf["k_bch"] // 8
            B = count[k]
            hdr = torch.from_numpy(bbheader(kb)).cuda()
            prbs = torch.from_numpy(ca.bb_prbs(kb)).cuda()
            pay = torch.randint(0, 256, (B, kb - 10), generator=g, device="cuda", dtype=torch.uint8)
            frame = torch.cat([hdr.expand(B, 10), pay], 1) ^ prbs
            sym = torch.empty((B, inf["xfec_len"], 2), dtype=torch.float32, device="cuda")
            cw = torch.empty((B, inf["n_ldpc"] // 8), dtype=torch.uint8, device="cuda")
            c.gen_frames_dev(seed * 131 + gi, B, es, sym, info_in=frame.contiguous(), cw_out=cw)
            if learned:
                path = learn.params_for(fs, rate, mod)
                if path:
                    _, bq, aq = learn.load_params(path)
                    c.set_params(bq, aq)
                    self.learned.add(gi)
                    self.params[gi] = (bq, aq)
            self.ctxs.append(c)
            sym_g.append(sym)
            pay_g.append(pay)
            cw_g.append(cw)
        # arrival order, offsets, per-frame constants
        slot = {k: 0 for k in self.keys}
        n = len(spec)
        self.n = n
        self.code = np.zeros(n, np.int32)
        self.row = np.zeros(n, np.int32)
        for f, s_ in enumerate(spec):
            self.code[f], self.row[f] = idx[s_], slot[s_]
            slot[s_] += 1
        info = [c.info for c in self.ctxs]
        self.N = np.array([info[c]["n_ldpc"] for c in self.code], np.int64)
        self.K = np.array([info[c]["k_ldpc"] for c in self.code], np.int64)
        self.kbch = np.array([info[c]["k_bch"] for c in self.code], np.int64)
        self.t_bch = np.array([info[c]["bch_t"] for c in self.code], np.int64)
        self.lt = np.array([info[c]["lt_alloc"] for c in self.code], np.int64)
        xf = np.array([info[c]["xfec_len"] for c in self.code], np.int64)
        self.sym_off = np.concatenate([[0], np.cumsum(xf)[:-1]])
        self.dlen = self.kbch // 8 - 10
        self.out_off = np.concatenate([[0], np.cumsum(self.dlen)[:-1]])
        self.llr_off = np.concatenate([[0], np.cumsum(self.N)[:-1]])
        self.hard_off = np.concatenate([[0], np.cumsum(self.K // 8)[:-1]])
        self.out_bytes = int(self.dlen.sum())
        self.sym = torch.empty((int(xf.sum()), 2), dtype=torch.float32, device="cuda")
        sent = torch.empty(self.out_bytes, dtype=torch.uint8, device="cuda")
        cw_all = torch.empty(int((self.K // 8).sum()), dtype=torch.uint8, device="cuda")

        def scatter(dst, src_rows, starts, width):
            # rows of one group into their arrival-order slots, one indexed copy
            idx = torch.from_numpy((starts[:, None] + np.arange(width)[None, :]).reshape(-1)).cuda()
            dst.index_copy_(0, idx, src_rows.reshape(idx.numel(), *dst.shape[1:]))

        for c in range(len(self.keys)):
            fr = np.nonzero(self.code == c)[0]
            fr = fr[np.argsort(self.row[fr])]  # group row order
            scatter(self.sym, sym_g[c], self.sym_off[fr], int(xf[fr[0]]))
            scatter(sent, pay_g[c], self.out_off[fr], int(self.dlen[fr[0]]))
            kb8 = int(self.K[fr[0]] // 8)
            scatter(cw_all, cw_g[c][:, :kb8].contiguous(), self.hard_off[fr], kb8)
        torch.cuda.synchronize()
        self.sent = sent.cpu().numpy()
        self.cw = cw_all.cpu().numpy()  # the BCH codewords (= systematic LDPC info part)
        self.recs = np.zeros(n, ca.STREAM_RX_FRAME)
        for f in range(n):
            nv, gain = snr_scales(self.keys[self.code[f]][3])
            self.recs[f] = (self.code[f], 0, self.sym_off[f], self.out_off[f], nv, gain)
        # learned iterations per frame (0: the code decodes plain)
        self.t_learned = np.array([self.params[c][0].shape[0] if c in self.params else 0 for c in self.code],
                                  np.int64)
        self.groups = len(self.keys)
        self.modcods = len(set((k[1], k[2]) for k in self.keys))
        self.short_fraction = float(np.mean([s_[0] == 1 for s_ in spec]))

    def score(self, trials, valid, dfl, out, hard=None):
        """Per-frame delivery (data field equal to the one sent) and, given
        the hard decisions (the work buffer, plan layout), the bit errors of
        the decoded BCH codeword against the sent one."""
        out = np.asarray(out)[:self.out_bytes]
        bad = np.add.reduceat(out != self.sent, self.out_off, dtype=np.int64)
        ok = (np.asarray(dfl) == self.dlen) & (bad == 0) & (np.asarray(trials) >= 0) & (np.asarray(valid) == 1)
        res = {"ok": ok}
        if hard is not None:
            x = _POPCOUNT[np.bitwise_xor(np.asarray(hard)[:self.cw.size], self.cw)]
            res["bit_errors"] = np.add.reduceat(x, self.hard_off, dtype=np.int64)
        return res


def run(workload="mixed12", n_frames=None, seed=20240612, reps=5, learned=None, log=print):
    import torch
    import cusat_amd as ca
    from cusat_amd import learn
    n_frames = n_frames or (2048 if workload == "mixed12" else 4096)
    learned = (workload == "s2x") if learned is None else learned
    spec = make_stream(workload, n_frames, seed)
    keys = sorted(set(spec))
    count = {k: sum(1 for s in spec if s == k) for k in keys}
    ctxs, llr_g, info_g = {}, {}, {}
    t0 = time.time()
    for gi, k in enumerate(keys):
        fs, rate, mod, es = k
        c = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=count[k])
        sym, info, _ = ca.gen_frames(fs, rate, mod, seed + 1000 * gi, count[k], es)
        nv, g = snr_scales(es)
        d = torch.empty((count[k], c.info["n_ldpc"]), dtype=torch.int16, device="cuda")
        c.demap(torch.from_numpy(sym).cuda(), d, count[k], nv, g)
        ctxs[k], llr_g[k], info_g[k] = c, d, info
    torch.cuda.synchronize()
    log("generated %d frames over %d (frame, MODCOD) groups in %.1f s" % (n_frames, len(keys), time.time() - t0))

    # learned parameters of every (frame size, MODCOD) that has a trained
    # file whose own evaluation found a gain
    have_learned = set()
    if learned:
        for k, c in ctxs.items():
            path = learn.params_for(k[0], k[1], k[2])
            if path:
                _, bq, aq = learn.load_params(path)
                c.set_params(bq, aq)
                have_learned.add(k)

    # pack the stream: frames interleaved in arrival order, LLRs back to back
    slot = {k: 0 for k in keys}
    order = []
    for s in spec:
        order.append((s, slot[s]))
        slot[s] += 1
    idx = {k: i for i, k in enumerate(keys)}
    frames = np.zeros(n_frames, ca.STREAM_FRAME)
    lo = ho = 0
    parts = []
    kb_total = 0
    for n, (k, j) in enumerate(order):
        c = ctxs[k]
        frames[n] = (idx[k], 0, lo, ho)
        parts.append(llr_g[k][j])
        lo += c.info["n_ldpc"]
        ho += c.info["k_ldpc"] // 8
        kb_total += c.info["k_bch"]
    llr = torch.cat(parts).contiguous()
    hard = torch.zeros(ho, dtype=torch.uint8, device="cuda")
    tr = torch.zeros(n_frames, dtype=torch.int32, device="cuda")
    va = torch.zeros(n_frames, dtype=torch.int32, device="cuda")
    fb = torch.zeros(n_frames, dtype=torch.int32, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    # baseline: one batched launch (decode + BCH) per (frame size, MODCOD) group
    ref = {}
    bufs = {}
    for k in keys:
        c = ctxs[k]
        B = count[k]
        bufs[k] = (torch.zeros((B, c.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda"),
                   torch.zeros(B, dtype=torch.int32, device="cuda"),
                   torch.zeros(B, dtype=torch.int32, device="cuda"),
                   torch.zeros(B, dtype=torch.int32, device="cuda"))

    def per_group(use_learned):
        for k in keys:
            c, B = ctxs[k], count[k]
            h, t, v, f = bufs[k]
            if use_learned and k in have_learned:
                c.rx_fec_learned(llr_g[k], h, t, v, f, B, 20)
            else:
                c.decode(llr_g[k], h, t, B)
                c.bch_check(h, v, B)

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(n):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / n

    ms_groups = timed(lambda: per_group(learned), reps)
    for k in keys:
        ref[k] = [x.cpu().numpy() for x in bufs[k]]

    sc = ca.StreamContext([ctxs[k] for k in keys], n_frames, ca.STREAM_LEARNED if learned else 0)
    sc.plan(frames)
    t1 = time.time()
    chosen, ms_tune = sc.autotune(llr, hard, tr, va, fb, reps=2)
    log("autotune: %s (%.3f ms per launch), %.1f s" % (chosen, ms_tune, time.time() - t1))
    ms_eager = timed(lambda: sc.decode(llr, hard, tr, va, fb), reps)

    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        sc.decode(llr, hard, tr, va, fb)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        sc.decode(llr, hard, tr, va, fb)
    ms_graph = timed(g.replay, reps)

    # frame-by-frame equality with the per-group launches
    hn, tn, vn, fn_ = hard.cpu().numpy(), tr.cpu().numpy(), va.cpu().numpy(), fb.cpu().numpy()
    same = True
    ok = np.zeros(n_frames, bool)
    iters = np.zeros(n_frames)
    for n, (k, j) in enumerate(order):
        h0, t0_, v0, f0 = ref[k]
        kb8 = ctxs[k].info["k_ldpc"] // 8
        o = frames[n]["hard_offset"]
        if tn[n] != t0_[j] or vn[n] != v0[j] or not np.array_equal(hn[o:o + kb8], h0[j]):
            same = False
        if learned and k in have_learned and fn_[n] != f0[j]:
            same = False
        kbb = ctxs[k].info["k_bch"] // 8
        ok[n] = tn[n] >= 0 and vn[n] == 1 and np.array_equal(hn[o:o + kbb], info_g[k][j])
        plain_it = (50 - tn[n]) if tn[n] >= 0 else 50
        if learned and k in have_learned:
            # learned pass of 20 iterations, then the plain re-decode
            iters[n] = 20 + plain_it if fn_[n] == 1 else 20 - tn[n]
        else:
            iters[n] = plain_it
    res = {
        "workload": workload, "frames": n_frames, "modcods": len(set((k[1], k[2]) for k in keys)),
        "groups": len(keys), "short_fraction": float(np.mean([s[0] == 1 for s in spec])),
        "learned_codes": len(have_learned), "decoder": "learned+fallback" if learned else "plain",
        "info_bits": int(kb_total),
        "ms_per_launch_graph": round(ms_graph, 4), "ms_per_launch_eager": round(ms_eager, 4),
        "ms_per_modcod_launches": round(ms_groups, 4),
        "gbps_info_stream": round(kb_total / (ms_graph * 1e-3) / 1e9, 3),
        "gbps_info_per_modcod": round(kb_total / (ms_groups * 1e-3) / 1e9, 3),
        "speedup": round(ms_groups / ms_graph, 3), "tuning": chosen,
        "fer": float(1 - ok.mean()), "mean_iterations": float(iters.mean()),
        "fallback_rate": float(fn_.mean()) if learned else 0.0,
        "identical_to_per_modcod": bool(same),
    }
    log(json.dumps(res))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mixed12", choices=["mixed12", "s2x"])
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--plain", action="store_true", help="s2x: no learned parameters")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    r = run(a.workload, a.frames, reps=a.reps, learned=False if a.plain else None,
            log=lambda s: print(s, flush=True))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(r, f, indent=1)
    return 0 if r["identical_to_per_modcod"] else 1


if __name__ == "__main__":
    sys.exit(main())
