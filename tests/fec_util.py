"""Shared helpers of the parity tests: run the HIP path through the C ABI and
the CPU oracle on the same seeded inputs."""
import numpy as np

import cusat_amd as ca
from oracle import oracle_ctypes as oc

TABLE = oc.TABLE

_codes = {}


def oracle_code(framesize, rate, use_sched=True):
    key = (TABLE[(framesize, rate)], use_sched)
    if key not in _codes:
        _codes[key] = oc.Code(key[0], use_sched)
    return _codes[key]


def snr_scales(es_n0_db):
    """noise_var and llr_gain exactly as rx_fec_chain::update_snr computes them
    (rx_fec_chain.h:148-156), in float32."""
    snr_lin = np.float32(10.0) ** (np.float32(es_n0_db) / np.float32(10.0))
    nv = np.float32(1.0) / np.float32(snr_lin)
    gain = min(np.float32(4.0) / nv, np.float32(1000.0))
    return float(nv), float(gain)


def dmax_by_definition(P, n_mod, mod):
    """demapper_kernel.cu:216-237 restated with numpy float32: QPSK 2, 8PSK
    4, APSK the largest, over (point, bit), of the smallest squared distance
    to a point with the other value of that bit."""
    if mod == "QPSK":
        return np.float32(2.0)
    if mod == "8PSK":
        return np.float32(4.0)
    P = np.asarray(P, np.float32).reshape(-1, 2)
    dr = (P[:, None, 0] - P[None, :, 0]).astype(np.float32)
    di = (P[:, None, 1] - P[None, :, 1]).astype(np.float32)
    d = (dr * dr + di * di).astype(np.float32)
    lab = np.arange(len(P))
    dmax = np.float32(0.0)
    for b in range(n_mod):
        m = 1 << (n_mod - 1 - b)
        opp = ((lab[:, None] ^ lab[None, :]) & m) != 0
        dmax = max(dmax, np.where(opp, d, np.float32(1e30)).min(1).max())
    return np.float32(dmax)


def reference_demap_scales(ctx, noise_var, llr_gain):
    """The demapper scaling as the reference computes it, from the context's
    constellation: QPSK 2 sqrt(2) / noise_var (demapper_kernel.cu:64); 8PSK
    2 sin(pi/8) times the capped gain (:84-99); the cap 2 * 320 / dmax
    (capped_llr_gain, :283-291)."""
    f32 = np.float32
    qs = f32(f32(2.0) * f32(1.4142135623730951)) / f32(noise_var)
    pts, _ = ctx.constellation()
    mod = ca.MODULATIONS[ctx.modulation]
    cap = f32(f32(2.0) * f32(320.0)) / dmax_by_definition(pts, ctx.info["n_mod"], mod)
    g = f32(llr_gain) if f32(llr_gain) < cap else cap
    ps = f32(f32(2.0) * f32(0.38268343236508977)) * g
    return float(qs), float(ps), float(g)


def oracle_demap(ctx, sym, noise_var, llr_gain):
    """The oracle's LLRs for the context's frames, with the demapper scaling
    computed independently of the library (and checked against what the
    library uses)."""
    qs, ps, g = reference_demap_scales(ctx, noise_var, llr_gain)
    got = ctx.demap_scales(noise_var, llr_gain)
    assert [np.float32(v) for v in got] == [np.float32(v) for v in (qs, ps, g)], (got, (qs, ps, g))
    pts, col = ctx.constellation()
    formula = ctx.modulation in (0, 1)
    return oc.demap(sym, ctx.info["n_mod"], None if formula else pts, col, ctx.info["n_ldpc"],
                    qs, ps, g, n_sym=ctx.info["demap_len"])


def gpu_chain(ctx, sym, noise_var, llr_gain):
    """demap -> decode -> BCH on the GPU through cusat_rx_fec_batch."""
    import torch
    B = sym.shape[0]
    inf = ctx.info
    dev = torch.device("cuda")
    sym_d = torch.from_numpy(np.ascontiguousarray(sym)).to(dev)
    llr_d = torch.zeros((B, inf["n_ldpc"]), dtype=torch.int16, device=dev)
    hard_d = torch.zeros((B, inf["k_ldpc"] // 8), dtype=torch.uint8, device=dev)
    tr_d = torch.full((B,), -7, dtype=torch.int32, device=dev)
    va_d = torch.full((B,), -7, dtype=torch.int32, device=dev)
    ctx.rx_fec_batch(sym_d, llr_d, B, noise_var, llr_gain, hard_d, tr_d, va_d)
    torch.cuda.synchronize()
    return llr_d.cpu().numpy(), hard_d.cpu().numpy(), tr_d.cpu().numpy(), va_d.cpu().numpy()


def gpu_decode(ctx, llr):
    import torch
    llr = np.ascontiguousarray(llr, np.int16)
    B = llr.shape[0]
    dev = torch.device("cuda")
    llr_d = torch.from_numpy(llr).to(dev)
    hard_d = torch.zeros((B, ctx.info["k_ldpc"] // 8), dtype=torch.uint8, device=dev)
    tr_d = torch.full((B,), -7, dtype=torch.int32, device=dev)
    ctx.decode(llr_d, hard_d, tr_d, B)
    torch.cuda.synchronize()
    return hard_d.cpu().numpy(), tr_d.cpu().numpy()


def gpu_bch(ctx, hard):
    import torch
    hard = np.ascontiguousarray(hard, np.uint8)
    B = hard.shape[0]
    dev = torch.device("cuda")
    h = torch.from_numpy(hard).to(dev)
    v = torch.full((B,), -7, dtype=torch.int32, device=dev)
    ctx.bch_check(h, v, B)
    torch.cuda.synchronize()
    return v.cpu().numpy()


def codeword_llr(cw_bytes, n, mag=127):
    bits = np.unpackbits(cw_bytes, axis=-1)[..., :n]
    return np.where(bits == 0, mag, -mag).astype(np.int16)


def dd_snr_host(sym, points, qpsk):
    """Decision-directed SNR as the host computes it (float32, symbol order)."""
    f32 = np.float32
    s = f32(0.0)
    e = f32(0.0)
    if qpsk:
        a = f32(0.70710678118654752)
        for re, im in sym.astype(np.float32):
            dr = -a if re < 0 else a
            di = -a if im < 0 else a
            s = f32(s + f32(f32(dr * dr) + f32(di * di)))
            e = f32(e + f32(f32(re - dr) * f32(re - dr)))
            e = f32(e + f32(f32(im - di) * f32(im - di)))
    else:
        pts = points.astype(np.float32)
        for re, im in sym[::8].astype(np.float32):
            dx = (re - pts[:, 0]).astype(np.float32)
            dy = (im - pts[:, 1]).astype(np.float32)
            d = (dx * dx + dy * dy).astype(np.float32)
            k = int(np.argmin(d))  # first minimum, as the host's strict < scan
            s = f32(s + f32(f32(pts[k, 0] * pts[k, 0]) + f32(pts[k, 1] * pts[k, 1])))
            e = f32(e + d[k])
    if not (np.isfinite(s) and np.isfinite(e)):
        return f32(0.0)
    return f32(1e6) if f32(e * f32(1e6)) <= s else f32(s / e)
