#!/usr/bin/env python3
"""Benchmark of the RX FEC hot path on one MI355X.

Headline workload: the combined DVB-S2/S2X ACM stream -- 4096 FECFRAMEs
drawn uniformly from 15 MODCODs (QPSK 1/4, 1/2, 2/3, 3/4; 8PSK 3/5, 3/4,
9/10; 16APSK 2/3, 3/4, 5/6; 32APSK 3/4, 9/10; 64APSK 4/5, 128APSK 3/4,
256APSK 3/4), 25 % short frames (DVB-S2 MODCODs only), each at its MODCOD's
threshold (lib/cusat_dsp/acm_config.h:54-121) + 0.3 dB.  Every frame is a
scrambled BBFRAME (valid BBHEADER, data field of seeded random bytes
filling the frame) sent through the GPU transmit chain (BCH, LDPC,
interleave + map, AWGN); the symbols sit in HBM before timing starts.

A step is one receive pass over the whole stream: the three launches of
cusat_stream_rx -- the mixed soft demap + deinterleave into int16 LLRs, the
persistent layered min-sum LDPC kernel with the BCH check fused (learned
offset / normalised min-sum with per-MODCOD trained parameters, every frame
failing the FEC checks re-decoded by the plain 50-iteration decoder in the
same workgroup), the BB descramble + data-field extraction.  `value` is
decoded info Gbit/s: k_bch bits of every frame whose delivered data field
equals the one sent, over the step time.  The all-plain decoder runs the
same stream in the same run ("plain").  FER is reported three ways: frames
not delivered; the reference bench's system FER (frames with more than t
bit errors in the decoded BCH codeword, apps/bench_coded_modulation.cc:
384-428); LDPC-alone FER (any bit error).

The JSON line carries the stream kernel's roofline (algorithmic bytes per
frame pass 2N + K/8 + 8 + I * 2 * LT_alloc, I = iterations executed, over
the decode launch's average duration from HIP events on the launch stream)
and the CPU baseline (the oracle's C restatement of the reference decoder,
learned passes and plain fallback composed the same way, on the LLRs the
GPU demapped).  "plain_worst_case" is DVB-S2 normal 8PSK 3/4, 1024 frames at
Es/N0 7.0 dB (below the waterfall: every frame runs 50 iterations) on the
batched decoder, all 1024 frames checked against the oracle;
"learned_vs_plain_16apsk" the learned decoder of DVB-S2 normal 16APSK 2/3
against the plain one at Es/N0 9.2 dB.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
T = 50


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=20240612)
    ap.add_argument("--cpu-reps", type=int, default=3, help="passes of the CPU baseline over the stream")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary workloads")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--worst-case", action="store_true",
                    help="run only the 8PSK 3/4 worst case on the batched decoder and print its JSON object")
    ap.add_argument("--no-split", action="store_true", help="worst case on the 6-wave decoder kernel")
    return ap.parse_args()


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_threads(avail):
    """Threads of the CPU baseline: every core this process may run on, up
    to the 16 host cores a one-GPU job is given."""
    return max(1, min(16, avail))


def stream_step(sc, sym, out, tr, va, fb, dfl, ev=None):
    if ev is not None:
        ev[0].record()
    sc.rx_demap(sym)
    if ev is not None:
        ev[1].record()
    sc.rx_decode(tr, va, fb)
    if ev is not None:
        ev[2].record()
    sc.rx_bb(out, tr, va, dfl)
    if ev is not None:
        ev[3].record()


def measure_stream(st, sc, steps, warmup, autotune):
    """Warm-up, then `steps` timed receive passes; per-phase HIP-event
    times, results of the last pass."""
    import torch
    import cusat_amd as ca
    n = st.n
    out = torch.zeros(st.out_bytes + 1, dtype=torch.uint8, device="cuda")
    tr, va, fb, dfl = (torch.zeros(n, dtype=torch.int32, device="cuda") for _ in range(4))
    sc.rx_plan_async(st.recs)
    tuning = None
    if autotune:
        sc.rx_demap(st.sym)
        a_llr, a_hard = sc.rx_work()
        tuning, _ = sc.autotune(a_llr, a_hard, tr, va, fb, reps=2)
        sc.rx_plan_async(st.recs)  # the receive plan under the chosen tuning
    for _ in range(warmup):
        stream_step(sc, st.sym, out, tr, va, fb, dfl)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        stream_step(sc, st.sym, out, tr, va, fb, dfl, evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ph = [float(np.mean([e[i].elapsed_time(e[i + 1]) for e in evs])) for i in range(3)]
    a_llr, a_hard = sc.rx_work()
    hard = ca.copy_from_device(a_hard, int((st.K // 8).sum()))
    res = {"ms_step": wall * 1e3 / steps, "ms_demap": ph[0], "ms_decode": ph[1], "ms_bb": ph[2],
           "trials": tr.cpu().numpy(), "valid": va.cpu().numpy(), "fallback": fb.cpu().numpy(),
           "dfl": dfl.cpu().numpy(), "out": out.cpu().numpy(), "hard": hard, "tuning": tuning or sc.tuning(),
           "llr_addr": a_llr}
    return res


def summarise(st, r, learned):
    sc = st.score(r["trials"], r["valid"], r["dfl"], r["out"], r["hard"])
    ok, be = sc["ok"], sc["bit_errors"]
    tr, fb = r["trials"], r["fallback"]
    plain_it = np.where(tr >= 0, T - tr, T).astype(np.float64)
    T_l = st.t_learned
    lmask = (T_l > 0) if learned else np.zeros(st.n, bool)
    # iterations executed: a learned frame's learned pass (T_l - trials, or
    # all T_l when the plain decoder redid it: an upper bound for a redone
    # frame that met its syndrome with a wrong codeword), plus the redo
    it_l = np.where(fb == 1, T_l + plain_it, np.where(tr >= 0, T_l - tr, T_l)).astype(np.float64)
    iters = np.where(lmask, it_l, plain_it)
    passes = np.where(lmask & (fb == 1), 2, 1)
    # algorithmic bytes of the decode launch
    alg = float(np.sum(passes * (2 * st.N + st.K // 8 + 8) + iters * 2 * st.lt))
    good_bits = float(np.sum(st.kbch[ok]))
    return {
        "gbps": good_bits / (r["ms_step"] * 1e-3) / 1e9,
        "gbps_all_frames": float(np.sum(st.kbch)) / (r["ms_step"] * 1e-3) / 1e9,
        "ms_per_step": r["ms_step"], "ms_demap": r["ms_demap"], "ms_decode": r["ms_decode"], "ms_bb": r["ms_bb"],
        "fer_delivered": float(1.0 - ok.mean()),
        "fer_system_ref_def": float(np.mean(be > st.t_bch)),
        "fer_ldpc": float(np.mean(be > 0)),
        "frames_delivered": int(ok.sum()),
        "mean_iterations": float(iters.mean()),
        "fallback_rate": float(fb.mean()) if learned else 0.0,
        "alg_bytes_decode": alg,
        "tuning": r["tuning"],
    }


def cpu_baseline(st, r, reps, nthr):
    """The oracle decodes every frame of the stream from the LLRs the GPU
    demapped: learned codes run their learned pass, then the plain decoder
    redoes the frames failing the FEC checks; other codes decode plain.
    Returns the timed rate and whether every frame equals the GPU's."""
    import cusat_amd as ca
    from oracle import oracle_ctypes as oc
    llr = ca.copy_from_device(r["llr_addr"], 2 * int(st.N.sum())).view(np.int16)
    groups = []
    for gi, k in enumerate(st.keys):
        fr = np.nonzero(st.code == gi)[0]
        L = np.stack([llr[st.llr_off[f]:st.llr_off[f] + st.N[f]] for f in fr])
        inf = st.ctxs[gi].info
        groups.append((gi, k, fr, L, oc.Code(oc.TABLE[(k[0], k[1])]), inf))
    params = st.params

    def decode_all():
        res = {}
        for gi, k, fr, L, code, inf in groups:
            if gi in params:
                bq, aq = params[gi]
                h, t = code.decode(L, bq.shape[0], beta=bq, alpha=aq, n_param_iters=bq.shape[0], nthreads=nthr)
                v = oc.bch_check(h, inf["n_bch"], k[0], inf["bch_t"])
                redo = np.nonzero(~((t >= 0) & (v == 1)))[0]
                if len(redo):
                    h2, t2 = code.decode(L[redo], T, nthreads=nthr)
                    h[redo], t[redo] = h2, t2
                    v[redo] = oc.bch_check(h2, inf["n_bch"], k[0], inf["bch_t"])
                fb = np.zeros(len(fr), np.int32)
                fb[redo] = 1
            else:
                h, t = code.decode(L, T, nthreads=nthr)
                v = oc.bch_check(h, inf["n_bch"], k[0], inf["bch_t"])
                fb = np.zeros(len(fr), np.int32)
            res[gi] = (h, t, v, fb)
        return res

    t0 = time.perf_counter()
    for _ in range(reps):
        res = decode_all()
    dt = time.perf_counter() - t0
    same = 0
    for gi, k, fr, L, code, inf in groups:
        h, t, v, fb = res[gi]
        for j, f in enumerate(fr):
            o, kb8 = st.hard_off[f], st.K[f] // 8
            same += int(t[j] == r["trials"][f] and v[j] == r["valid"][f] and fb[j] == r["fallback"][f] and
                        np.array_equal(h[j], r["hard"][o:o + kb8]))
    return dt, same


def plain_worst_case(nthr, split=True, reps=5):
    """DVB-S2 normal 8PSK 3/4, 1024 frames at Es/N0 7.0 dB, T = 50 (every
    frame at the iteration cap), demap -> batched decode -> BCH; all 1024
    frames against the oracle."""
    import torch
    import cusat_amd as ca
    from oracle import oracle_ctypes as oc
    fs, rate, mod, B, es = 0, "3/4", "8PSK", 1024, 7.0
    ctx = ca.FecContext(fs, rate, mod, max_trials=T, max_batch=B, flags=0 if split else ca.CTX_NO_SPLIT_KERNEL)
    inf = ctx.info
    N, K, kb = inf["n_ldpc"], inf["k_ldpc"], inf["k_bch"]
    sym, _, _ = ca.gen_frames(fs, rate, mod, 20240607, B, es)
    from cusat_amd.stream_bench import snr_scales
    nv, gain = snr_scales(es)
    sym_d = torch.from_numpy(sym).cuda()
    llr = torch.empty((B, N), dtype=torch.int16, device="cuda")
    hard = torch.empty((B, K // 8), dtype=torch.uint8, device="cuda")
    tr, va = (torch.empty(B, dtype=torch.int32, device="cuda") for _ in range(2))
    for _ in range(2):
        ctx.rx_fec_batch(sym_d, llr, B, nv, gain, hard, tr, va)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(reps)]
    t0 = time.perf_counter()
    for e in ev:
        e[0].record()
        ctx.demap(sym_d, llr, B, nv, gain)
        e[1].record()
        ctx.decode(llr, hard, tr, B)
        e[2].record()
        ctx.bch_check(hard, va, B)
        e[3].record()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    ms_dec = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    t = tr.cpu().numpy()
    h = hard.cpu().numpy()
    iters = np.where(t >= 0, T - t, T)
    alg = float(np.sum(2 * N + K // 8 + 8 + iters.astype(np.float64) * 2 * inf["lt_alloc"]))
    h_o, t_o = oc.Code("s2_b7").decode(llr.cpu().numpy(), T, nthreads=nthr)
    traffic = None
    tj = os.path.join(ROOT, "profiles", "ldpc_kernel_traffic.json")
    if os.path.exists(tj) and split:
        traffic = json.load(open(tj)).get("hbm_bytes_per_launch")
    return {"workload": "DVB-S2 normal 8PSK 3/4 plain, 1024 frames at 7.0 dB, demap + batched decode + BCH",
            "kernel": "ldpc_kernel_split<12, true, 2>" if split else "ldpc_kernel<12, true, 2> (6-wave)",
            "ms_bch": round(float(np.mean([e[2].elapsed_time(e[3]) for e in ev])), 4),
            "ms_demap": round(float(np.mean([e[0].elapsed_time(e[1]) for e in ev])), 4),
            "gbps_all_frames": round(B * kb / (ms * 1e-3) / 1e9, 4), "ms_per_step": round(ms, 4),
            "ms_ldpc": round(ms_dec, 4), "mean_iterations": float(iters.mean()),
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms_dec * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / (ms_dec * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": traffic},
            "oracle_parity_frames": B,
            "oracle_parity": bool(np.array_equal(h_o, h) and np.array_equal(t_o, t))}


def learned_vs_plain_16apsk(frames=16384, es=9.2):
    """DVB-S2 normal 16APSK 2/3: learned decoder with plain fallback
    against the plain decoder, each one persistent launch over `frames`
    fresh GPU frames at Es/N0 `es`."""
    from cusat_amd import learn
    rows = learn.evaluate_stream(points=(es,), frames=frames, source="gpu", reps=3, log=lambda s: None)
    r = {x["decoder"].split(" ")[0]: x for x in rows}
    keep = ("gbps_info", "fer", "mean_iterations", "ms_launch")
    res = {"workload": "DVB-S2 normal 16APSK 2/3, fresh GPU-generated frames, one persistent launch each",
           "es_n0_db": es, "frames": frames,
           "plain": {k: round(r["plain"][k], 6) for k in keep},
           "learned_fallback": {k: round(r["learned+fallback"][k], 6) for k in keep},
           "identical_to_batched": bool(r["plain"]["identical_to_batched"] and
                                        r["learned+fallback"]["identical_to_batched"])}
    res["speedup"] = round(res["learned_fallback"]["gbps_info"] / res["plain"]["gbps_info"], 4)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards by frame index with no exchange; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import torch
    import cusat_amd as ca
    from cusat_amd import stream_bench

    if args.worst_case:
        print(json.dumps(plain_worst_case(cpu_threads(cpu_info()[2]), split=not args.no_split,
                                          reps=args.steps)), flush=True)
        return
    st = stream_bench.RxStream("s2x", args.frames, args.seed, learned=True, max_trials=T)
    sc_l = ca.StreamContext(st.ctxs, st.n, ca.STREAM_LEARNED)
    sc_p = ca.StreamContext(st.ctxs, st.n, 0)
    r_l = measure_stream(st, sc_l, args.steps, args.warmup, not args.no_autotune)
    r_p = measure_stream(st, sc_p, args.steps, args.warmup, not args.no_autotune)
    s_l = summarise(st, r_l, True)
    s_p = summarise(st, r_p, False)
    achieved = s_l["alg_bytes_decode"] / (s_l["ms_decode"] * 1e-3) / 1e9

    model, ncpu, avail = cpu_info()
    nthr = cpu_threads(avail)
    cpu = None
    parity = None
    if not args.no_cpu:
        dt, same = cpu_baseline(st, r_l, args.cpu_reps, nthr)
        good = float(np.sum(st.kbch[st.score(r_l["trials"], r_l["valid"], r_l["dfl"], r_l["out"])["ok"]]))
        cpu = {"value": round(args.cpu_reps * good / dt / 1e9, 6), "unit": "Gbit/s", "cores": nthr,
               "kind": "port", "cpu_model": model, "nproc": ncpu,
               "sample": "the whole %d-frame stream, %d passes, decode + BCH only (LLRs from the GPU demapper): "
                         "oracle C decoder (-O3), learned passes + plain redo of failing frames as the GPU, "
                         "%d threads, %.1f s" % (st.n, args.cpu_reps, nthr, dt)}
        parity = {"frames": st.n, "identical": same}

    traffic = None
    tj = os.path.join(ROOT, "profiles", "ldpc_stream_s2x_traffic.json")
    if os.path.exists(tj) and (args.frames, args.seed) == (4096, 20240612):
        traffic = json.load(open(tj)).get("hbm_bytes_per_launch")

    def rnd(d):
        return {k: (round(v, 6) if isinstance(v, float) else v) for k, v in d.items() if k != "alg_bytes_decode"}

    out = {
        "metric": "decoded info Gbit/s at FER 1e-3, learned vs plain min-sum, 1 MI355X",
        "value": round(s_l["gbps"], 4),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_l["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic (seeded BBFRAMEs -> BCH + LDPC -> interleave/map -> AWGN on the GPU)",
        "config": {"workload": "DVB-S2/S2X combined ACM stream: %d frames over %d MODCODs (%d codes), "
                               "%.0f%% short, threshold + 0.3 dB; demap -> persistent LDPC + BCH "
                               "(learned + plain fallback) -> BB" % (st.n, st.modcods, st.groups,
                                                                     100 * st.short_fraction),
                   "frames_per_step": st.n, "max_trials": T,
                   "learned_trials": sorted(set(int(x) for x in st.t_learned if x)),
                   "learned_codes": len(st.learned), "decoder": "learned + fallback",
                   "info_bits": "k_bch of every frame whose data field is delivered intact",
                   "speedup_vs_plain": round(s_l["gbps"] / s_p["gbps"], 4),
                   "learned": rnd(s_l), "plain": rnd(s_p),
                   "oracle_parity": parity},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "ldpc_stream_kernel (learned + fallback, BCH fused)"},
        "cpu_baseline": cpu,
    }
    if not args.no_extra:
        out["plain_worst_case"] = plain_worst_case(nthr)
        out["learned_vs_plain_16apsk"] = learned_vs_plain_16apsk()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
