"""Learned offset / normalised min-sum: training and evaluation driver.

Training (fp32, HIP forward/backward kernels of cusat_learn_fwd_bwd):
per-iteration x per-layer offsets beta[t][i] and scales alpha[t][i] of the
check node, over n_iters unrolled layered iterations, BCE of every
iteration's APP against the transmitted codeword.  Data: seeded random
codewords from the host transmit chain (BCH -> IRA LDPC -> interleave ->
map), AWGN added on the device at an Es/N0 drawn per sub-batch, and the
int16 LLRs of the product demapper (cusat_fec_demap) -- the plain pipeline.

Inference quantises the parameters to the integer decoder's hooks:
beta -> round, clamped to 0..255 (LLR units; plain = 1); alpha -> n/64 with
n = round(64 alpha) clamped to 1..1023 (the decoder's odd-symmetric
(|v| n + 32) >> 6 scale, lib/cusat_dsp/cuda/ldpc_decode_kernel.cu:75-90).

    python -m cusat_amd.learn train --out cusat_amd/data/learned/s2_b6_16apsk.json
    python -m cusat_amd.learn eval  --params cusat_amd/data/learned/s2_b6_16apsk.json
    python -m cusat_amd.learn train-set --codes s2x     # one file per MODCOD and frame size
"""
import argparse
import glob
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LEARNED_DIR = os.path.join(HERE, "data", "learned")
DEFAULT_PARAMS = os.path.join(LEARNED_DIR, "s2_b6_16apsk.json")

# MODCODs of the mixed ACM streams with their quasi-error-free Es/N0
# thresholds in dB (lib/cusat_dsp/acm_config.h:54-121): the twelve DVB-S2
# MODCODs of the 12-MODCOD stream, then the three DVB-S2X ones.  No short
# FECFRAME exists for rate 9/10 or for the S2X APSK codes used here.
MODCODS12 = [("1/4", "QPSK", -0.05), ("1/2", "QPSK", 1.80), ("3/4", "QPSK", 4.43),
             ("3/5", "8PSK", 5.80), ("9/10", "8PSK", 11.08), ("2/3", "16APSK", 9.17),
             ("5/6", "16APSK", 11.71), ("3/4", "32APSK", 12.93), ("9/10", "32APSK", 16.05),
             ("2/3", "QPSK", 3.50), ("3/4", "8PSK", 8.11), ("3/4", "16APSK", 10.41)]
S2X = [("4/5", "64APSK", 16.10), ("135/180", "128APSK", 18.00), ("135/180", "256APSK", 23.34)]
NO_SHORT = {"9/10", "135/180"}
NO_SHORT_MOD = {"64APSK", "128APSK", "256APSK"}


def has_short(rate, mod):
    return rate not in NO_SHORT and mod not in NO_SHORT_MOD


def params_name(fs, rate, mod):
    return "%s_%s_%s.json" % ("normal" if fs == 0 else "short", mod.lower(), rate.replace("/", "-"))


def _code_of(path):
    with open(path) as fh:
        d = json.load(fh)
    return d.get("framesize"), d.get("rate"), d.get("modulation")


def params_for(fs, rate, mod, directory=LEARNED_DIR):
    """Path of the trained parameters of (frame size, rate, modulation) in
    `directory`, or None; a file whose own evaluation found no gain
    ("use": false) is skipped."""
    for f in sorted(glob.glob(os.path.join(directory, "*.json"))):
        with open(f) as fh:
            d = json.load(fh)
        if (d.get("framesize"), d.get("rate"), d.get("modulation")) == (fs, rate, mod) and d.get("use", True):
            return f
    return None


def quantize(alpha, beta):
    """fp32 (alpha, beta) -> int16 (beta, alpha64) for cusat_ldpc_set_params."""
    b = np.clip(np.rint(np.asarray(beta, np.float64)), 0, 255).astype(np.int16)
    a = np.clip(np.rint(64.0 * np.asarray(alpha, np.float64)), 1, 1023).astype(np.int16)
    return b, a


def load_params(path=DEFAULT_PARAMS):
    d = json.load(open(path))
    return d, np.array(d["beta_q"], np.int16), np.array(d["alpha_q"], np.int16)


def snr_scales(es_n0_db):
    """noise_var, llr_gain as rx_fec_chain::update_snr (rx_fec_chain.h:148-156)."""
    snr_lin = np.float32(10.0) ** (np.float32(es_n0_db) / np.float32(10.0))
    nv = np.float32(1.0) / snr_lin
    return float(nv), float(min(np.float32(4.0) / nv, np.float32(1000.0)))


class FramePool:
    """Noiseless XFECFRAMEs of seeded random codewords (host generator),
    resident on the device, plus their codeword bits; noisy LLR batches are
    drawn from it."""

    def __init__(self, ctx, framesize, rate, modulation, n_frames, seed):
        import torch
        import cusat_amd as ca
        self.ctx = ctx
        sym, info, cw = ca.gen_frames(framesize, rate, modulation, seed, n_frames, 1000.0, want_cw=True)
        self.n = n_frames
        self.sym = torch.from_numpy(sym).cuda()
        self.cw = torch.from_numpy(cw).cuda()
        self.info = torch.from_numpy(info).cuda()
        self.N = ctx.info["n_ldpc"]

    def llr(self, idx, es_db, gen, llr_out, sub=32):
        """int16 LLRs [len(idx), N] into llr_out; es_db is one value per
        sub-batch of `sub` frames (one noise level per demap call)."""
        import torch
        B = len(idx)
        for k0 in range(0, B, sub):
            k1 = min(B, k0 + sub)
            es = float(es_db[k0 // sub])
            nv, gain = snr_scales(es)
            s = self.sym[idx[k0:k1]]
            noise = torch.randn(s.shape, generator=gen, device=s.device) * float(np.sqrt(nv / 2.0))
            noisy = (s + noise).contiguous()
            self.ctx.demap(noisy, llr_out[k0:k1], k1 - k0, nv, gain)

    def draw(self, idx, es_db, gen, llr_out, sub=32):
        """Noisy frames of pool rows idx -> llr_out; returns their (info
        bytes, codeword bytes) on the device."""
        self.llr(idx, es_db, gen, llr_out, sub)
        return self.info[idx], self.cw[idx]


class GpuFrames:
    """Fresh frames for every draw from the GPU transmit chain
    (cusat_gen_frames_dev): new random codewords, AWGN at one Es/N0 per
    sub-batch, demapped into llr_out -- no host generator in the loop."""

    def __init__(self, ctx, max_batch, seed):
        import torch
        self.ctx = ctx
        g = ctx.info
        self.seed = int(seed)
        self.calls = 0
        self.sym = torch.empty((max_batch, g["xfec_len"], 2), dtype=torch.float32, device="cuda")
        self.info = torch.empty((max_batch, g["k_bch"] // 8), dtype=torch.uint8, device="cuda")
        self.cw = torch.empty((max_batch, g["n_ldpc"] // 8), dtype=torch.uint8, device="cuda")

    def draw(self, B, es_db, llr_out, sub=32):
        for k0 in range(0, B, sub):
            k1 = min(B, k0 + sub)
            es = float(es_db[k0 // sub])
            nv, gain = snr_scales(es)
            self.calls += 1
            self.ctx.gen_frames_dev(self.seed * 1000003 + self.calls, k1 - k0, es, self.sym[k0:k1],
                                    info_out=self.info[k0:k1], cw_out=self.cw[k0:k1])
            self.ctx.demap(self.sym[k0:k1], llr_out[k0:k1], k1 - k0, nv, gain)
        return self.info[:B], self.cw[:B]


def train(framesize=0, rate="2/3", modulation="16APSK", n_iters=20, batch=512, steps=400,
          es_lo=8.9, es_hi=9.6, pool_frames=2048, seed=20240611, kappa=0.5, lr_alpha=0.005,
          lr_beta=0.02, quantized=True, log=print, frames="host", alpha0=1.0, beta0=1.0):
    """frames: "host" = a pool of host-generated codewords re-noised per
    batch; "gpu" = fresh codewords per batch from the GPU generator.
    alpha0, beta0: the starting scale and offset of every (iteration,
    layer) (1, 1 = the plain decoder)."""
    import torch
    import cusat_amd as ca
    ctx = ca.FecContext(framesize, rate, modulation, max_trials=50, max_batch=max(batch, 64))
    q, N = ctx.info["q"], ctx.info["n_ldpc"]
    if frames == "gpu":
        pool = GpuFrames(ctx, batch, seed)
    else:
        pool = FramePool(ctx, framesize, rate, modulation, pool_frames, seed)
    lc = ca.LearnContext(ctx, batch, n_iters)
    dev = torch.device("cuda")
    alpha = torch.full((n_iters, q), float(alpha0), dtype=torch.float32, device=dev)
    beta = torch.full((n_iters, q), float(beta0), dtype=torch.float32, device=dev)
    ma, va, mb, vb = [torch.zeros_like(alpha) for _ in range(4)]
    ga, gb = torch.zeros_like(alpha), torch.zeros_like(beta)
    loss = torch.zeros(n_iters, dtype=torch.float32, device=dev)
    llr16 = torch.empty((batch, N), dtype=torch.int16, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for step in range(1, steps + 1):
        if frames == "gpu":
            es = rng.uniform(es_lo, es_hi, (batch + 31) // 32)
            _, cw = pool.draw(batch, es, llr16)
        else:
            idx = torch.from_numpy(rng.choice(pool.n, batch, replace=False)).to(dev)
            es = rng.uniform(es_lo, es_hi, (batch + 31) // 32)
            _, cw = pool.draw(idx, es, gen, llr16)
        llrf = llr16.float()
        cw = cw.contiguous()
        ga.zero_(), gb.zero_(), loss.zero_()
        lc.fwd_bwd(llrf, cw, batch, alpha, beta, kappa, ga, gb, loss, quantized=quantized)
        ca.adam_step(alpha, ga, ma, va, lr_alpha, step, 0.25, 2.0)
        ca.adam_step(beta, gb, mb, vb, lr_beta, step, 0.0, 16.0)
        if step == 1 or step % 10 == 0 or step == steps:
            l = loss.cpu().numpy()
            hist.append({"step": step, "loss_mean": float(l.mean()), "loss_last": float(l[-1]),
                         "alpha_mean": float(alpha.mean()), "beta_mean": float(beta.mean()),
                         "seconds": round(time.time() - t0, 1)})
            log("train step %d: loss %.5f (last iteration %.5f) alpha %.3f beta %.3f  %.1fs" %
                (step, l.mean(), l[-1], float(alpha.mean()), float(beta.mean()), time.time() - t0))
    a = alpha.cpu().numpy()
    b = beta.cpu().numpy()
    bq, aq = quantize(a, b)
    lc.close()
    return {"framesize": framesize, "rate": rate, "modulation": modulation, "q": q,
            "n_iters": n_iters, "alpha": a.round(5).tolist(), "beta": b.round(5).tolist(),
            "alpha_q": aq.tolist(), "beta_q": bq.tolist(),
            "training": {"batch": batch, "steps": steps, "es_n0_db": [es_lo, es_hi],
                         "pool_frames": pool_frames if frames == "host" else None, "frames": frames,
                         "seed": seed, "kappa": kappa,
                         "lr_alpha": lr_alpha, "lr_beta": lr_beta, "quantized": quantized,
                         "alpha0": alpha0, "beta0": beta0,
                         "history": hist}}


def evaluate(params_path=DEFAULT_PARAMS, points=(8.8, 9.0, 9.2, 9.4, 9.6), frames=16384, chunk=1024,
             learned_trials=20, plain_trials=50, pool_frames=2048, seed=777, log=print, params=None,
             source="host"):
    """FER and decoded info Gbit/s per Es/N0 of: the plain decoder (T = 50),
    the learned decoder alone (T = learned_trials) and the learned decoder
    with plain fallback, on the same noisy LLRs.  `params`: a parameter dict
    as train() returns, instead of the file.  source: "host" (re-noised
    pool of host-generated codewords) or "gpu" (fresh GPU-generated frames)."""
    import torch
    import cusat_amd as ca
    if params is None:
        d, bq, aq = load_params(params_path)
    else:
        d = params
        bq, aq = np.array(d["beta_q"], np.int16), np.array(d["alpha_q"], np.int16)
    fs, rate, mod = d["framesize"], d["rate"], d["modulation"]
    chunk = min(chunk, frames)
    ctx = ca.FecContext(fs, rate, mod, max_trials=plain_trials, max_batch=chunk)
    ctx.set_params(bq, aq)
    # the same structure (learned_trials + plain fallback) with the plain
    # check node, to separate the parameters' gain from the early cut-off's
    ctx_id = ca.FecContext(fs, rate, mod, max_trials=plain_trials, max_batch=chunk)
    ctx_id.set_params(np.ones_like(bq), np.full_like(aq, 64))
    inf = ctx.info
    N, K, kb = inf["n_ldpc"], inf["k_ldpc"], inf["k_bch"]
    pool = GpuFrames(ctx, chunk, seed) if source == "gpu" else FramePool(ctx, fs, rate, mod, pool_frames, seed)
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    llr = torch.empty((chunk, N), dtype=torch.int16, device=dev)
    hard = torch.empty((chunk, K // 8), dtype=torch.uint8, device=dev)
    tr = torch.empty(chunk, dtype=torch.int32, device=dev)
    va = torch.empty(chunk, dtype=torch.int32, device=dev)
    fb = torch.empty(chunk, dtype=torch.int32, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    rows = []
    for es in points:
        acc = {m: {"err": 0, "ldpc_fail": 0, "iters": 0, "ms": 0.0, "fallback": 0}
               for m in ("plain", "learned", "learned+fallback", "identity+fallback")}
        for c0 in range(0, frames, chunk):
            es_sub = np.full((chunk + 31) // 32, es)
            if source == "gpu":
                info_ref, _ = pool.draw(chunk, es_sub, llr)
            else:
                info_ref, _ = pool.draw(torch.from_numpy(rng.integers(0, pool.n, chunk)).to(dev), es_sub, gen, llr)
            info_ref = info_ref.clone()
            for m in acc:
                torch.cuda.synchronize()
                ev[0].record()
                if m == "plain":
                    ctx.decode(llr, hard, tr, chunk)
                    ctx.bch_check(hard, va, chunk)
                elif m == "learned":
                    ctx.decode_learned(llr, hard, tr, chunk, learned_trials)
                    ctx.bch_check(hard, va, chunk)
                elif m == "learned+fallback":
                    ctx.rx_fec_learned(llr, hard, tr, va, fb, chunk, learned_trials)
                else:
                    ctx_id.rx_fec_learned(llr, hard, tr, va, fb, chunk, learned_trials)
                ev[1].record()
                torch.cuda.synchronize()
                a = acc[m]
                a["ms"] += ev[0].elapsed_time(ev[1])
                ok = (tr >= 0) & (va == 1)
                good = ok & (hard[:, : kb // 8] == info_ref).all(1)
                a["err"] += int((~good).sum())
                a["ldpc_fail"] += int((tr < 0).sum())
                T = learned_trials if m == "learned" else plain_trials
                it = torch.where(tr >= 0, T - tr, torch.full_like(tr, T))
                if m.endswith("+fallback"):
                    a["fallback"] += int(fb.sum())
                    it = torch.where(fb == 1, it + learned_trials, learned_trials - tr)
                a["iters"] += int(it.sum())
        for m, a in acc.items():
            row = {"es_n0_db": es, "decoder": m, "frames": frames, "fer": a["err"] / frames,
                   "ldpc_fail_rate": a["ldpc_fail"] / frames, "mean_iterations": a["iters"] / frames,
                   "gbps_info": frames * kb / (a["ms"] * 1e-3) / 1e9, "ms_total": round(a["ms"], 3)}
            if m.endswith("+fallback"):
                row["fallback_rate"] = a["fallback"] / frames
            rows.append(row)
            log(json.dumps(row))
    return rows


def evaluate_stream(params_path=DEFAULT_PARAMS, points=(9.0, 9.2, 9.4), frames=32768, learned_trials=20,
                    plain_trials=50, pool_frames=2048, seed=777, reps=3, log=print, source="host"):
    """The plain decoder vs the learned decoder with plain fallback as a
    receiver runs them on a continuous stream: all `frames` frames of a
    point (LLRs resident in HBM) decoded by ONE persistent launch
    (cusat_stream_decode: decode, BCH and, in learned mode, the plain
    re-decode of failing frames inside the same workgroup), timed with HIP
    events as the mean of `reps` launches.  Frame-by-frame results are
    compared with the batched decoders (cusat_ldpc_decode + cusat_bch_check,
    cusat_rx_fec_learned) on the same LLRs."""
    import torch
    import cusat_amd as ca
    d, bq, aq = load_params(params_path)
    fs, rate, mod = d["framesize"], d["rate"], d["modulation"]
    chunk = min(1024, frames)
    ctx = ca.FecContext(fs, rate, mod, max_trials=plain_trials, max_batch=frames)
    ctx.set_params(bq, aq)
    inf = ctx.info
    N, K, kb = inf["n_ldpc"], inf["k_ldpc"], inf["k_bch"]
    pool = GpuFrames(ctx, chunk, seed) if source == "gpu" else FramePool(ctx, fs, rate, mod, pool_frames, seed)
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    llr = torch.empty((frames, N), dtype=torch.int16, device=dev)
    hard = torch.empty((frames, K // 8), dtype=torch.uint8, device=dev)
    tr = torch.empty(frames, dtype=torch.int32, device=dev)
    va = torch.empty(frames, dtype=torch.int32, device=dev)
    fb = torch.empty(frames, dtype=torch.int32, device=dev)
    plan = np.zeros(frames, ca.STREAM_FRAME)
    plan["llr_offset"] = np.arange(frames, dtype=np.int64) * N
    plan["hard_offset"] = np.arange(frames, dtype=np.int64) * (K // 8)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    streams = {m: ca.StreamContext([ctx], frames, f)
               for m, f in (("plain", 0), ("learned+fallback", ca.STREAM_LEARNED))}
    for sc in streams.values():
        sc.plan(plan)
    rows = []
    for es in points:
        info_ref = torch.empty((frames, kb // 8), dtype=torch.uint8, device=dev)
        es_sub = np.full((chunk + 31) // 32, es)
        idx = None if source == "gpu" else torch.from_numpy(rng.integers(0, pool.n, frames)).to(dev)
        for c0 in range(0, frames, chunk):
            if source == "gpu":
                inf, _ = pool.draw(chunk, es_sub, llr[c0:c0 + chunk])
            else:
                inf, _ = pool.draw(idx[c0:c0 + chunk], es_sub, gen, llr[c0:c0 + chunk])
            info_ref[c0:c0 + chunk] = inf
        for m, sc in streams.items():
            sc.decode(llr, hard, tr, va, fb)
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(reps):
                sc.decode(llr, hard, tr, va, fb)
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / reps
            # the batched decoders on the same LLRs, chunk by chunk
            same = True
            h2 = torch.empty((chunk, K // 8), dtype=torch.uint8, device=dev)
            t2, v2, f2 = [torch.empty(chunk, dtype=torch.int32, device=dev) for _ in range(3)]
            for c0 in range(0, frames, chunk):
                sl = slice(c0, c0 + chunk)
                if m == "plain":
                    ctx.decode(llr[sl], h2, t2, chunk)
                    ctx.bch_check(h2, v2, chunk)
                else:
                    ctx.rx_fec_learned(llr[sl], h2, t2, v2, f2, chunk, learned_trials)
                    same &= bool(torch.equal(f2, fb[sl]))
                same &= bool(torch.equal(h2, hard[sl]) and torch.equal(t2, tr[sl]) and torch.equal(v2, va[sl]))
            ok = (tr >= 0) & (va == 1) & (hard[:, : kb // 8] == info_ref).all(1)
            it = torch.where(tr >= 0, plain_trials - tr, torch.full_like(tr, plain_trials))
            if m != "plain":
                it = torch.where(fb == 1, it + learned_trials, learned_trials - tr)
            row = {"es_n0_db": es, "decoder": m + " (stream)", "frames": frames,
                   "fer": float((~ok).float().mean()), "mean_iterations": float(it.float().mean()),
                   "gbps_info": frames * kb / (ms * 1e-3) / 1e9, "ms_launch": round(ms, 3),
                   "identical_to_batched": same}
            if m != "plain":
                row["fallback_rate"] = float(fb.float().mean())
            rows.append(row)
            log(json.dumps(row))
    for sc in streams.values():
        sc.close()
    return rows


def train_set(codes="s2x", steps=200, batch=512, pool_frames=1024, out_dir=LEARNED_DIR, margin=0.3,
              eval_frames=2048, force=False, only=None, log=print, code=None, lr_alpha=0.005, lr_beta=0.02,
              alpha0=1.0, beta0=1.0):
    """Train one parameter file per (frame size, MODCOD) of a mixed stream
    (codes: "mixed12" or "s2x"), each over Es/N0 in [threshold - 0.2,
    threshold + margin + 0.2], then evaluate it at threshold + margin
    against the plain decoder on the same LLRs.  only: modulations to
    train; code: one (frame size, rate, modulation) to train; lr_alpha,
    lr_beta, alpha0, beta0 as in train().  The file records the
    evaluation and "use": the learned decoder with plain fallback decoded
    no more frames wrongly than the plain one, in fewer mean iterations."""
    mods = MODCODS12 + (S2X if codes == "s2x" else [])
    if only:
        mods = [m for m in mods if m[1] in only]
    os.makedirs(out_dir, exist_ok=True)
    done = []
    for rate, mod, th in mods:
        for fs in ((0, 1) if has_short(rate, mod) else (0,)):
            if code is not None and (fs, rate, mod) != tuple(code):
                continue
            path = os.path.join(out_dir, params_name(fs, rate, mod))
            have = [f for f in glob.glob(os.path.join(out_dir, "*.json"))
                    if _code_of(f) == (fs, rate, mod)]
            if have and not force:
                log("skip %s %s %s: %s" % (fs, mod, rate, os.path.basename(have[0])))
                continue
            t0 = time.time()
            log("train %s %s %s: Es/N0 %.2f..%.2f dB" % ("normal" if fs == 0 else "short", mod, rate,
                                                         th - 0.2, th + margin + 0.2))
            d = train(fs, rate, mod, batch=batch, steps=steps, es_lo=th - 0.2, es_hi=th + margin + 0.2,
                      pool_frames=pool_frames, log=lambda s: None, lr_alpha=lr_alpha, lr_beta=lr_beta,
                      alpha0=alpha0, beta0=beta0)
            rows = evaluate(points=(th + margin,), frames=eval_frames, pool_frames=pool_frames,
                            log=lambda s: None, params=d)
            r = {x["decoder"]: x for x in rows}
            p, lf = r["plain"], r["learned+fallback"]
            d["eval"] = {"es_n0_db": th + margin, "frames": eval_frames,
                         "plain": {k: p[k] for k in ("fer", "mean_iterations", "gbps_info")},
                         "learned+fallback": {k: lf[k] for k in ("fer", "mean_iterations", "gbps_info",
                                                                  "fallback_rate")}}
            d["use"] = bool(lf["fer"] <= p["fer"] and lf["mean_iterations"] < p["mean_iterations"])
            with open(path, "w") as f:
                json.dump(d, f, indent=1)
            log("  %.1f s: plain FER %.2e iters %.2f | learned+fallback FER %.2e iters %.2f fallback %.3f"
                " -> use %s" % (time.time() - t0, p["fer"], p["mean_iterations"], lf["fer"],
                                lf["mean_iterations"], lf["fallback_rate"], d["use"]))
            done.append(path)
    return done


def fer_crossing(rows, decoder, target=1e-3):
    """Es/N0 where the FER of `decoder` falls through `target` (log-linear
    interpolation between the sweep points) and the Gbit/s there."""
    r = sorted([x for x in rows if x["decoder"] == decoder], key=lambda x: x["es_n0_db"])
    for a, b in zip(r, r[1:]):
        if a["fer"] > target >= b["fer"]:
            fa, fb = np.log10(max(a["fer"], 1e-9)), np.log10(max(b["fer"], 1e-9))
            w = (fa - np.log10(target)) / (fa - fb) if fa != fb else 0.0
            return (a["es_n0_db"] + w * (b["es_n0_db"] - a["es_n0_db"]),
                    a["gbps_info"] + w * (b["gbps_info"] - a["gbps_info"]))
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=DEFAULT_PARAMS)
    t.add_argument("--steps", type=int, default=400)
    t.add_argument("--batch", type=int, default=512)
    t.add_argument("--iters", type=int, default=20)
    t.add_argument("--pool", type=int, default=2048)
    t.add_argument("--lr-alpha", type=float, default=0.005)
    t.add_argument("--lr-beta", type=float, default=0.02)
    t.add_argument("--float-model", action="store_true", help="train the fp32 model without rounding")
    t.add_argument("--es", default="8.9,9.6", help="training Es/N0 range in dB")
    t.add_argument("--gen", default="host", choices=["host", "gpu"],
                   help="training frames: re-noised host-generated pool, or fresh GPU-generated frames")
    ts = sub.add_parser("train-set")
    ts.add_argument("--codes", default="s2x", choices=["mixed12", "s2x"])
    ts.add_argument("--steps", type=int, default=200)
    ts.add_argument("--batch", type=int, default=512)
    ts.add_argument("--pool", type=int, default=1024)
    ts.add_argument("--eval-frames", type=int, default=2048)
    ts.add_argument("--out-dir", default=LEARNED_DIR)
    ts.add_argument("--only", default=None, help="comma-separated modulations to train")
    ts.add_argument("--force", action="store_true", help="retrain codes that have a file")
    ts.add_argument("--code", default=None, help="one code to train: normal|short,RATE,MODULATION")
    ts.add_argument("--lr-alpha", type=float, default=0.005)
    ts.add_argument("--lr-beta", type=float, default=0.02)
    ts.add_argument("--alpha0", type=float, default=1.0, help="starting scale of every (iteration, layer)")
    ts.add_argument("--beta0", type=float, default=1.0, help="starting offset of every (iteration, layer)")
    e = sub.add_parser("eval")
    e.add_argument("--params", default=DEFAULT_PARAMS)
    e.add_argument("--frames", type=int, default=16384)
    e.add_argument("--points", default="8.8,9.0,9.2,9.4,9.6")
    e.add_argument("--out", default=None)
    e.add_argument("--stream", action="store_true",
                   help="one persistent launch per point over all frames (plain vs learned + fallback)")
    e.add_argument("--gen", default="host", choices=["host", "gpu"],
                   help="frames: re-noised pool of host-generated codewords, or fresh GPU-generated ones")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    if a.cmd == "train":
        lo, hi = [float(x) for x in a.es.split(",")]
        d = train(n_iters=a.iters, batch=a.batch, steps=a.steps, pool_frames=a.pool, es_lo=lo, es_hi=hi,
                  lr_alpha=a.lr_alpha, lr_beta=a.lr_beta, quantized=not a.float_model, log=log,
                  frames=a.gen)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            json.dump(d, f, indent=1)
        log("wrote " + a.out)
    elif a.cmd == "train-set":
        code = None
        if a.code:
            fs, rate, mod = a.code.split(",")
            code = ({"normal": 0, "short": 1}[fs], rate, mod)
        train_set(a.codes, a.steps, a.batch, a.pool, a.out_dir, eval_frames=a.eval_frames, force=a.force,
                  only=a.only.split(",") if a.only else None, log=log, code=code, lr_alpha=a.lr_alpha,
                  lr_beta=a.lr_beta, alpha0=a.alpha0, beta0=a.beta0)
    else:
        pts = [float(x) for x in a.points.split(",")]
        if a.stream:
            rows = evaluate_stream(a.params, pts, a.frames, log=log, source=a.gen)
        else:
            rows = evaluate(a.params, pts, a.frames, log=log, source=a.gen)
        summary = {m: fer_crossing(rows, m) for m in sorted(set(r["decoder"] for r in rows))}
        log(json.dumps({"fer_1e-3_crossing": summary}))
        if a.out:
            with open(a.out, "w") as f:
                json.dump({"rows": rows, "fer_1e-3_crossing": summary}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
