"""GPU test-traffic generator (cusat_gen_frames_dev) against the host
transmit chain (cusat_encode_frames, itself pinned by the noiseless decoder
round trips and the BCH / LDPC checks of the oracle):

* the same info bytes give bit-identical codewords and noiseless symbols,
  for PSK / APSK normal and short frames, S2X codes, the 128APSK padding
  symbols and an explicit-coordinate 256APSK constellation;
* random info bytes are reproducible by seed, differ between seeds, and
  encode exactly as the host encodes them;
* the AWGN is zero-mean with variance N0/2 per component, uncorrelated
  between I and Q, Gaussian (kurtosis 3);
* frames generated on the GPU above threshold decode to their info bytes.
"""
import numpy as np
import pytest

import cusat_amd as ca
from tests.fec_util import snr_scales

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _gen(ctx, seed, B, es, info=None, want=True):
    torch = _torch()
    g = ctx.info
    sym = torch.full((B, g["xfec_len"], 2), 7.0, dtype=torch.float32, device="cuda")
    inf = torch.zeros((B, g["k_bch"] // 8), dtype=torch.uint8, device="cuda") if want else None
    cw = torch.zeros((B, g["n_ldpc"] // 8), dtype=torch.uint8, device="cuda") if want else None
    d_info = torch.from_numpy(np.ascontiguousarray(info)).cuda() if info is not None else None
    ctx.gen_frames_dev(seed, B, es, sym, info_in=d_info, info_out=inf, cw_out=cw)
    torch.cuda.synchronize()
    return (sym.cpu().numpy(), inf.cpu().numpy() if want else None, cw.cpu().numpy() if want else None)


@pytest.mark.parametrize("fs,rate,mod", [(0, "1/2", "QPSK"), (0, "3/4", "8PSK"), (1, "2/3", "16APSK"),
                                         (0, "9/10", "32APSK"), (0, "135/180", "128APSK"),
                                         (0, "20/30", "256APSK"), (1, "7/15", "QPSK"),
                                         (0, "1/4", "QPSK")])
def test_given_info_matches_host_encoder(fs, rate, mod):
    rng = np.random.default_rng(3)
    g = ca.geometry(fs, rate, mod)
    B = 5
    info = rng.integers(0, 256, (B, g["k_bch"] // 8), dtype=np.uint8)
    info[0] = 0
    info[1] = 0xff
    ctx = ca.FecContext(fs, rate, mod, max_trials=10, max_batch=B)
    sym, inf, cw = _gen(ctx, 1, B, 1000.0, info=info)
    sym_h, cw_h = ca.encode_frames(fs, rate, mod, 1, info, 1000.0, want_cw=True)
    np.testing.assert_array_equal(inf, info)
    np.testing.assert_array_equal(cw, cw_h)
    np.testing.assert_array_equal(sym.view(np.uint32), sym_h.view(np.uint32))


def test_random_info_reproducible_and_encoded_as_host():
    fs, rate, mod, B = 0, "2/3", "16APSK", 6
    ctx = ca.FecContext(fs, rate, mod, max_trials=10, max_batch=B)
    s1, i1, c1 = _gen(ctx, 77, B, 1000.0)
    s2, i2, c2 = _gen(ctx, 77, B, 1000.0)
    s3, i3, _ = _gen(ctx, 78, B, 1000.0)
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_array_equal(s1, s2)
    assert (i1 != i3).mean() > 0.99 * 255 / 256 - 0.01
    assert len({bytes(r) for r in i1}) == B  # frames differ
    # bytes look uniform: each of the 8 bit positions about half ones
    bits = np.unpackbits(i1, axis=1)
    assert abs(bits.mean() - 0.5) < 0.005
    sym_h, cw_h = ca.encode_frames(fs, rate, mod, 0, i1, 1000.0, want_cw=True)
    np.testing.assert_array_equal(c1, cw_h)
    np.testing.assert_array_equal(s1.view(np.uint32), sym_h.view(np.uint32))


def test_awgn_statistics():
    fs, rate, mod, B, es = 0, "3/4", "8PSK", 16, 8.0
    ctx = ca.FecContext(fs, rate, mod, max_trials=10, max_batch=B)
    noisy, info, _ = _gen(ctx, 5, B, es)
    clean, _, _ = _gen(ctx, 5, B, 1000.0, info=info, want=False)
    n = (noisy - clean).astype(np.float64)
    var = 10.0 ** (-es / 10.0) / 2.0
    x = n.reshape(-1, 2)
    m = x.shape[0]
    assert np.all(np.abs(x.mean(0)) < 5 * np.sqrt(var / m))
    assert np.all(np.abs(x.var(0) / var - 1.0) < 0.01)
    assert abs(np.corrcoef(x[:, 0], x[:, 1])[0, 1]) < 0.01
    kurt = (x ** 4).mean(0) / x.var(0) ** 2
    assert np.all(np.abs(kurt - 3.0) < 0.05)
    # neighbouring symbols are uncorrelated too
    assert abs(np.corrcoef(x[:-1, 0], x[1:, 0])[0, 1]) < 0.01


@pytest.mark.parametrize("fs,rate,mod,es", [(0, "3/4", "8PSK", 9.0), (1, "1/2", "QPSK", 3.0)])
def test_gpu_traffic_decodes(fs, rate, mod, es):
    torch = _torch()
    B = 24
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=B)
    g = ctx.info
    sym = torch.empty((B, g["xfec_len"], 2), dtype=torch.float32, device="cuda")
    info = torch.empty((B, g["k_bch"] // 8), dtype=torch.uint8, device="cuda")
    ctx.gen_frames_dev(123, B, es, sym, info_out=info)
    nv, gain = snr_scales(es)
    llr = torch.empty((B, g["n_ldpc"]), dtype=torch.int16, device="cuda")
    hard = torch.empty((B, g["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
    tr = torch.empty(B, dtype=torch.int32, device="cuda")
    va = torch.empty(B, dtype=torch.int32, device="cuda")
    ctx.rx_fec_batch(sym, llr, B, nv, gain, hard, tr, va)
    torch.cuda.synchronize()
    assert (tr >= 0).all() and (va == 1).all()
    assert torch.equal(hard[:, : g["k_bch"] // 8], info)
