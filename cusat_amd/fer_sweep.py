"""Frame-error-rate studies on the GPU (measurement, not the product path).

thresholds: the decoder against the reference's published quality table.
The min_snr_db column of lib/cusat_dsp/acm_config.h:41-121 is, per
MODCOD, the first waterfall point (0.1 dB grid) whose post-BCH FER is at
most 1e-5, measured with normal FECFRAMEs, 25 LDPC iterations and AWGN by
apps/bench_coded_modulation.cc.  That bench counts a frame as an LDPC error
when its decoded BCH codeword differs from the sent one in any bit, and as
a system (post-BCH) error when more than t bits differ (the outer BCH code
corrects up to t, :384-428); it runs until 2M frames or 200 errors.  Here:
the same definitions, the GPU transmit chain (random info -> BCH -> LDPC
-> map -> AWGN), the demapper scaling of rx_fec_chain::update_snr, the
batched plain decoder with T = 25.  Es/N0 rises in 0.1 dB steps from the
published value - 0.3 dB; a point fails as soon as it holds more than 20
system errors (its FER over 2M frames would exceed 1e-5) and passes after
2M frames with at most 20; the scan stops at the first passing point.

crossing: learned offset/normalised min-sum (20 learned iterations, plain
50-iteration fallback in the same launch) against the plain decoder (50
iterations) on DVB-S2 normal 16APSK 2/3: FER per Es/N0 point over >= 1M
fresh frames with 95 % Wilson intervals, and the Es/N0 where each FER
crosses 1e-3 (log-linear interpolation between grid points; the interval
from interpolating the Wilson bounds).  FER = frames not accepted
(frame_passes_fec_checks, rx_fec_chain.h:193-196) or accepted with wrong
info bits; the reference-bench system FER (> t bit errors) beside it.

    python -m cusat_amd.fer_sweep thresholds --out thresholds.json
    python -m cusat_amd.fer_sweep crossing --out crossing.json
"""
import argparse
import json
import math
import sys
import time

import numpy as np

# (rate, modulation, published Es/N0 threshold) of the 12 MODCODs of the
# mixed ACM stream, acm_config.h:54-108
MODCODS12 = [("1/4", "QPSK", -0.05), ("1/2", "QPSK", 1.80), ("2/3", "QPSK", 3.50), ("3/4", "QPSK", 4.43),
             ("3/5", "8PSK", 5.80), ("3/4", "8PSK", 8.11), ("9/10", "8PSK", 11.08), ("2/3", "16APSK", 9.17),
             ("3/4", "16APSK", 10.41), ("5/6", "16APSK", 11.71), ("3/4", "32APSK", 12.93),
             ("9/10", "32APSK", 16.05)]


def snr_scales(es_n0_db):
    s = np.float32(10.0) ** (np.float32(es_n0_db) / np.float32(10.0))
    nv = np.float32(1.0) / s
    return float(nv), float(min(np.float32(4.0) / nv, np.float32(1000.0)))


def wilson(k, n, z=1.959964):
    if n == 0:
        return 0.0, 1.0
    p = k / n
    den = 1 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return max(0.0, c - h), min(1.0, c + h)


class Chain:
    """GPU transmit chain + demap for one MODCOD, batches of B frames."""

    def __init__(self, fs, rate, mod, T, B, seed):
        import torch
        import cusat_amd as ca
        self.torch = torch
        self.ctx = ca.FecContext(fs, rate, mod, max_trials=T, max_batch=B)
        inf = self.ctx.info
        self.inf, self.B, self.seed = inf, B, seed
        self.sym = torch.empty((B, inf["xfec_len"], 2), dtype=torch.float32, device="cuda")
        self.cw = torch.empty((B, inf["n_ldpc"] // 8), dtype=torch.uint8, device="cuda")
        self.info = torch.empty((B, inf["k_bch"] // 8), dtype=torch.uint8, device="cuda")
        self.llr = torch.empty((B, inf["n_ldpc"]), dtype=torch.int16, device="cuda")
        self.hard = torch.empty((B, inf["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
        self.tr = torch.empty(B, dtype=torch.int32, device="cuda")
        self.va = torch.empty(B, dtype=torch.int32, device="cuda")
        self.fb = torch.empty(B, dtype=torch.int32, device="cuda")
        self.pop = torch.tensor([bin(v).count("1") for v in range(256)], dtype=torch.int32, device="cuda")
        self.batches = 0

    def draw(self, es):
        """A fresh batch at Es/N0 es: LLRs in self.llr, sent codewords and info."""
        self.batches += 1
        self.ctx.gen_frames_dev(self.seed * 1000003 + self.batches, self.B, es, self.sym, info_out=self.info,
                                cw_out=self.cw)
        nv, g = snr_scales(es)
        self.ctx.demap(self.sym, self.llr, self.B, nv, g)

    def bit_errors(self):
        """Bit errors of the decoded BCH codeword (the K info bits of the
        LDPC code) per frame."""
        kb8 = self.inf["k_ldpc"] // 8
        x = (self.hard ^ self.cw[:, :kb8]).long()
        return self.pop[x].sum(1)


def thresholds(modcods=MODCODS12, T=25, B=8192, max_frames=2_000_000, max_sys_errors=20, start=-0.3, step=0.1,
               seed=11, log=print):
    import torch
    out = []
    for rate, mod, th in modcods:
        ch = Chain(0, rate, mod, T, B, seed)
        t_bch = ch.inf["bch_t"]
        kb = ch.inf["k_bch"] // 8
        found = None
        k = 0
        points = []
        t0 = time.time()
        while found is None and k < 12:
            es = round(th + start + k * step, 2)
            fr = ldpc_e = sys_e = chain_e = 0
            it = 0.0
            while fr < max_frames and sys_e <= max_sys_errors:
                ch.draw(es)
                ch.ctx.decode(ch.llr, ch.hard, ch.tr, ch.B)
                ch.ctx.bch_check(ch.hard, ch.va, ch.B)
                be = ch.bit_errors()
                acc = (ch.tr >= 0) & (ch.va == 1) & (ch.hard[:, :kb] == ch.info).all(1)
                fr += ch.B
                ldpc_e += int((be > 0).sum())
                sys_e += int((be > t_bch).sum())
                chain_e += int((~acc).sum())
                it += float(torch.where(ch.tr >= 0, T - ch.tr, torch.full_like(ch.tr, T)).float().sum())
            passed = sys_e <= max_sys_errors
            row = {"es_n0_db": es, "frames": fr, "ldpc_errors": ldpc_e, "system_errors": sys_e,
                   "chain_errors": chain_e, "fer_ldpc": ldpc_e / fr, "fer_system": sys_e / fr,
                   "fer_chain": chain_e / fr, "mean_iterations": it / fr, "passes_1e-5": passed}
            points.append(row)
            log(json.dumps({"modcod": "%s %s" % (mod, rate), **row}))
            if passed:
                found = es
            k += 1
        res = {"modcod": "%s %s" % (mod, rate), "published_es_n0_db": th, "first_pass_es_n0_db": found,
               "gap_db": None if found is None else round(found - th, 2), "seconds": round(time.time() - t0, 1),
               "points": points}
        out.append(res)
        log(json.dumps({k_: v for k_, v in res.items() if k_ != "points"}))
        del ch
        torch.cuda.empty_cache()
    return out


def crossing(points=(8.9, 9.0, 9.1, 9.2, 9.3, 9.4), frames=1_048_576, B=16384, seed=23, log=print):
    """Learned + fallback vs plain on DVB-S2 normal 16APSK 2/3, one
    persistent stream launch per batch and decoder on the same LLRs."""
    import torch
    import cusat_amd as ca
    from cusat_amd import learn
    d, bq, aq = learn.load_params(learn.DEFAULT_PARAMS)
    fs, rate, mod = d["framesize"], d["rate"], d["modulation"]
    ch = Chain(fs, rate, mod, 50, B, seed)
    ch.ctx.set_params(bq, aq)
    inf = ch.inf
    N, K, kb = inf["n_ldpc"], inf["k_ldpc"], inf["k_bch"] // 8
    t_bch = inf["bch_t"]
    plan = np.zeros(B, ca.STREAM_FRAME)
    plan["llr_offset"] = np.arange(B, dtype=np.int64) * N
    plan["hard_offset"] = np.arange(B, dtype=np.int64) * (K // 8)
    scs = {"plain": ca.StreamContext([ch.ctx], B, 0), "learned+fallback": ca.StreamContext([ch.ctx], B,
                                                                                            ca.STREAM_LEARNED)}
    for sc in scs.values():
        sc.plan(plan)
    rows = {m: [] for m in scs}
    for es in points:
        acc = {m: {"frames": 0, "chain_errors": 0, "system_errors": 0, "iterations": 0.0, "fallback": 0}
               for m in scs}
        while acc["plain"]["frames"] < frames:
            ch.draw(es)
            for m, sc in scs.items():
                sc.decode(ch.llr, ch.hard, ch.tr, ch.va, ch.fb)
                be = ch.bit_errors()
                ok = (ch.tr >= 0) & (ch.va == 1) & (ch.hard[:, :kb] == ch.info).all(1)
                a = acc[m]
                a["frames"] += B
                a["chain_errors"] += int((~ok).sum())
                a["system_errors"] += int((be > t_bch).sum())
                pl = torch.where(ch.tr >= 0, 50 - ch.tr, torch.full_like(ch.tr, 50))
                if m != "plain":
                    pl = torch.where(ch.fb == 1, pl + 20, 20 - ch.tr)
                    a["fallback"] += int(ch.fb.sum())
                a["iterations"] += float(pl.float().sum())
        for m, a in acc.items():
            lo, hi = wilson(a["chain_errors"], a["frames"])
            row = {"es_n0_db": es, "decoder": m, "frames": a["frames"], "errors": a["chain_errors"],
                   "fer": a["chain_errors"] / a["frames"], "fer_lo": lo, "fer_hi": hi,
                   "fer_system_ref_def": a["system_errors"] / a["frames"],
                   "mean_iterations": a["iterations"] / a["frames"], "fallback_rate": a["fallback"] / a["frames"]}
            rows[m].append(row)
            log(json.dumps(row))
    res = {"workload": "DVB-S2 normal 16APSK 2/3, learned (20 it.) + plain fallback vs plain (50 it.)",
           "points": rows, "crossing_1e-3": {}}
    for m, rs in rows.items():
        res["crossing_1e-3"][m] = {"es_n0_db": interp_crossing(rs, "fer"),
                                   "lo": interp_crossing(rs, "fer_lo"), "hi": interp_crossing(rs, "fer_hi")}
    c = res["crossing_1e-3"]
    if c["plain"]["es_n0_db"] is not None and c["learned+fallback"]["es_n0_db"] is not None:
        res["learned_minus_plain_db"] = round(c["learned+fallback"]["es_n0_db"] - c["plain"]["es_n0_db"], 4)
    log(json.dumps({k: v for k, v in res.items() if k != "points"}))
    for sc in scs.values():
        sc.close()
    return res


def interp_crossing(rows, key, target=1e-3):
    """Es/N0 where log10(key) crosses target, linear between grid points."""
    for a, b in zip(rows, rows[1:]):
        fa, fb = a[key], b[key]
        if fa >= target > fb and fb > 0:
            la, lb, lt = math.log10(fa), math.log10(fb), math.log10(target)
            return round(a["es_n0_db"] + (la - lt) / (la - lb) * (b["es_n0_db"] - a["es_n0_db"]), 4)
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["thresholds", "crossing"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--only", default=None, help="thresholds: comma-separated 'MOD rate' names")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.study == "thresholds":
        mods = MODCODS12
        if a.only:
            want = set(x.strip() for x in a.only.split(","))
            mods = [m for m in mods if "%s %s" % (m[1], m[0]) in want]
        r = thresholds(mods, max_frames=a.frames or 2_000_000, log=log)
    else:
        r = crossing(frames=a.frames or 1_048_576, log=log)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(r, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
