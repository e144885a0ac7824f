"""Shared helpers of the parity tests: run the HIP path through the C ABI and
the CPU oracle on the same seeded inputs."""
import numpy as np

import cusat_amd as ca
from oracle import oracle_ctypes as oc

TABLE = {  # (framesize, rate) -> oracle table name
    (0, "1/4"): "s2_b1", (0, "1/3"): "s2_b2", (0, "2/5"): "s2_b3", (0, "1/2"): "s2_b4",
    (0, "3/5"): "s2_b5", (0, "2/3"): "s2_b6", (0, "3/4"): "s2_b7", (0, "4/5"): "s2_b8",
    (0, "5/6"): "s2_b9", (0, "8/9"): "s2_b10", (0, "9/10"): "s2_b11",
    (0, "135/180"): "s2x_b19", (0, "128/180"): "s2x_b17", (0, "154/180"): "s2x_b21",
    (0, "96/180"): "s2x_b12", (0, "22/30"): "s2x_b24", (0, "7/9"): "s2x_b10", (1, "7/15"): "s2x_c4",
    (1, "1/4"): "s2_c1", (1, "1/3"): "s2_c2", (1, "2/5"): "s2_c3", (1, "1/2"): "s2_c4",
    (1, "3/5"): "s2_c5", (1, "2/3"): "s2_c6", (1, "3/4"): "s2_c7", (1, "4/5"): "s2_c8",
    (1, "5/6"): "s2_c9", (1, "8/9"): "s2_c10",
    # the other DVB-S2X codes, by their table in EN 302 307-2 annexes B and C
    (0, "13/45"): "s2x_b2", (0, "9/20"): "s2x_b3", (0, "11/20"): "s2x_b4", (0, "26/45"): "s2x_b5",
    (0, "28/45"): "s2x_b6", (0, "23/36"): "s2x_b7", (0, "25/36"): "s2x_b8", (0, "13/18"): "s2x_b9",
    (0, "90/180"): "s2x_b11", (0, "100/180"): "s2x_b13", (0, "104/180"): "s2x_b14",
    (0, "116/180"): "s2x_b15", (0, "124/180"): "s2x_b16", (0, "132/180"): "s2x_b18",
    (0, "140/180"): "s2x_b20", (0, "18/30"): "s2x_b22", (0, "20/30"): "s2x_b23",
    (1, "11/45"): "s2x_c1", (1, "4/15"): "s2x_c2", (1, "14/45"): "s2x_c3", (1, "8/15"): "s2x_c5",
    (1, "26/45"): "s2x_c6", (1, "32/45"): "s2x_c7",
}

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


def oracle_demap(ctx, sym, noise_var, llr_gain):
    qs, ps, g = ctx.demap_scales(noise_var, llr_gain)
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
