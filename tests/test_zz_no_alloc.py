"""cusat_ldpc_decode_noalloc (message records in the library's static device
scratch) against cusat_ldpc_decode (records in the context's workspace),
through the C ABI: two calls in a row on the same GPU, the second with a
smaller batch, both bit-exact (hard bits and trials); the batch is bounded
by the scratch, not by the context's max_batch; an oversize batch is
rejected on the host with CUSAT_E_BATCH before anything is launched."""
import ctypes

import numpy as np
import pytest

import cusat_amd as ca
from tests.fec_util import snr_scales

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _llrs(fs, rate, mod, seed, B, es):
    torch = _torch()
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=B)
    sym, _, _ = ca.gen_frames(fs, rate, mod, seed, B, es)
    nv, g = snr_scales(es)
    d = torch.empty((B, ctx.info["n_ldpc"]), dtype=torch.int16, device="cuda")
    ctx.demap(torch.from_numpy(sym).cuda(), d, B, nv, g)
    torch.cuda.synchronize()
    return ctx, d


def _decode(fn, ctx, llr, B):
    torch = _torch()
    hard = torch.full((B, ctx.info["k_ldpc"] // 8), 0xA5, dtype=torch.uint8, device="cuda")
    tr = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(ctx._h, ctypes.c_void_p(llr.data_ptr()), ctypes.c_void_p(hard.data_ptr()),
            ctypes.c_void_p(tr.data_ptr()), B, s)
    assert rc == 0, ca.lib().cusat_last_error().decode()
    torch.cuda.synchronize()
    return hard.cpu().numpy(), tr.cpu().numpy()


# (framesize, rate, modulation, Es/N0, first batch, second batch): near the
# waterfall, so that both converging and failing frames occur
CASES = [(0, "3/4", "8PSK", 8.0, 40, 7), (1, "1/2", "QPSK", 1.9, 33, 5), (0, "1/4", "QPSK", -0.2, 24, 3)]


@pytest.mark.parametrize("fs,rate,mod,es,b1,b2", CASES)
def test_noalloc_matches_decode_twice(fs, rate, mod, es, b1, b2):
    L = ca.lib()
    ref_ctx, llr = _llrs(fs, rate, mod, 77, b1, es)
    # the no-alloc call's context only needs its tables: a workspace for one
    # frame, while the call decodes b1 frames
    small = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=1)
    h0, t0 = _decode(L.cusat_ldpc_decode, ref_ctx, llr, b1)
    h1, t1 = _decode(L.cusat_ldpc_decode_noalloc, small, llr, b1)
    np.testing.assert_array_equal(t1, t0)
    np.testing.assert_array_equal(h1, h0)
    # second call, smaller and on other frames: the scratch keeps the first
    # call's records, which must not leak into the result
    sub = llr[b1 - b2:].contiguous()
    h2, t2 = _decode(L.cusat_ldpc_decode_noalloc, small, sub, b2)
    np.testing.assert_array_equal(t2, t0[b1 - b2:])
    np.testing.assert_array_equal(h2, h0[b1 - b2:])
    h3, t3 = _decode(L.cusat_ldpc_decode, ref_ctx, sub, b2)
    np.testing.assert_array_equal(t2, t3)
    np.testing.assert_array_equal(h2, h3)


def test_noalloc_rejects_oversize_batch():
    torch = _torch()
    L = ca.lib()
    ctx = ca.FecContext(0, "3/4", "8PSK", max_trials=50, max_batch=1)
    llr = torch.zeros((1, 64800), dtype=torch.int16, device="cuda")
    hard = torch.zeros((1, 6075), dtype=torch.uint8, device="cuda")
    tr = torch.zeros(1, dtype=torch.int32, device="cuda")
    # 8PSK 3/4: 259,200 bytes of message records per frame
    too_many = SCRATCH_BYTES // 259200 + 1
    for bad in (0, too_many, 1 << 20):
        rc = L.cusat_ldpc_decode_noalloc(ctx._h, ctypes.c_void_p(llr.data_ptr()), ctypes.c_void_p(hard.data_ptr()),
                                         ctypes.c_void_p(tr.data_ptr()), bad, None)
        assert rc == -4, (bad, rc)
        assert "static scratch" in L.cusat_last_error().decode()
    torch.cuda.synchronize()
