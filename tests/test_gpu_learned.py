"""Learned offset/normalised min-sum on the GPU, through the C ABI.

* Integer inference (cusat_ldpc_decode_learned) is bit-exact against the
  oracle with the same per-(iteration, layer) parameters, and identity
  parameters reproduce the plain decoder.
* cusat_rx_fec_learned = learned pass, BCH, plain re-decode of the frames
  that fail frame_passes_fec_checks: compared with that composition of
  oracle calls.
* The fp32 training kernels (cusat_learn_fwd_bwd) against the torch fp64
  reference of the same model: final APP, per-iteration losses and the
  gradients of alpha, beta and the input LLRs.  Tolerances are written in
  each assertion (fp32 kernels with atomically summed updates vs fp64).
"""
import numpy as np
import pytest

import cusat_amd as ca
from oracle import oracle_ctypes as oc
from tests.fec_util import oracle_code, oracle_demap, snr_scales

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _llr_frames(ctx, fs, rate, mod, seed, B, es):
    sym, info, cw = ca.gen_frames(fs, rate, mod, seed, B, es, want_cw=True)
    nv, g = snr_scales(es)
    return oracle_demap(ctx, sym, nv, g), info, cw


def _learned_gpu(ctx, llr, T):
    torch = _torch()
    B = llr.shape[0]
    d = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    hard = torch.zeros((B, ctx.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
    tr = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    ctx.decode_learned(d, hard, tr, B, T)
    torch.cuda.synchronize()
    return hard.cpu().numpy(), tr.cpu().numpy()


def _random_params(q, T, seed):
    rng = np.random.default_rng(seed)
    beta = rng.integers(0, 4, (T, q)).astype(np.int16)
    alpha = rng.integers(40, 90, (T, q)).astype(np.int16)
    return beta, alpha


@pytest.mark.parametrize("fs,rate,mod,es,B", [(0, "2/3", "16APSK", 9.0, 16), (1, "1/2", "QPSK", 1.6, 24)])
def test_identity_parameters_equal_plain(fs, rate, mod, es, B):
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=B)
    llr, _, _ = _llr_frames(ctx, fs, rate, mod, 31, B, es)
    q = ctx.info["q"]
    ctx.set_params(np.ones((20, q), np.int16), np.full((20, q), 64, np.int16))
    h1, t1 = _learned_gpu(ctx, llr, 50)
    h0, t0 = oracle_code(fs, rate).decode(llr, 50)
    np.testing.assert_array_equal(h1, h0)
    np.testing.assert_array_equal(t1, t0)


@pytest.mark.parametrize("fs,rate,mod,es,B,T", [(0, "2/3", "16APSK", 9.0, 16, 20),
                                                (0, "2/3", "16APSK", 8.6, 8, 30),
                                                (1, "1/2", "QPSK", 1.6, 24, 20),
                                                (0, "3/4", "8PSK", 7.8, 8, 20)])
def test_learned_decode_vs_oracle(fs, rate, mod, es, B, T):
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=B)
    llr, _, _ = _llr_frames(ctx, fs, rate, mod, 5, B, es)
    q = ctx.info["q"]
    beta, alpha = _random_params(q, 20, 9)
    ctx.set_params(beta, alpha)
    h1, t1 = _learned_gpu(ctx, llr, T)
    h0, t0 = oracle_code(fs, rate).decode(llr, T, beta=beta, alpha=alpha, n_param_iters=20)
    np.testing.assert_array_equal(t1, t0)
    np.testing.assert_array_equal(h1, h0)


def test_rx_fec_learned_fallback_vs_oracle():
    torch = _torch()
    # 14 learned iterations at 9.1 dB with scales near 1: about half of the
    # frames converge (checked with the oracle), the rest are redone
    fs, rate, mod, B, Tl = 0, "2/3", "16APSK", 32, 14
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=B)
    llr, info, _ = _llr_frames(ctx, fs, rate, mod, 12, B, 9.1)
    q = ctx.info["q"]
    rng = np.random.default_rng(4)
    beta = np.ones((Tl, q), np.int16)
    alpha = rng.integers(60, 69, (Tl, q)).astype(np.int16)
    ctx.set_params(beta, alpha)
    code = oracle_code(fs, rate)
    h_l, t_l = code.decode(llr, Tl, beta=beta, alpha=alpha, n_param_iters=Tl)
    v_l = oc.bch_check(h_l, ctx.info["n_bch"], fs, ctx.info["bch_t"])
    redo = ~((t_l >= 0) & (v_l == 1))
    assert redo.any() and (~redo).any(), "the case needs both kinds of frames"
    h_p, t_p = code.decode(llr, 50)
    v_p = oc.bch_check(h_p, ctx.info["n_bch"], fs, ctx.info["bch_t"])
    d = torch.from_numpy(llr).cuda()
    hard = torch.zeros((B, ctx.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
    tr = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    va = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    fb = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    ctx.rx_fec_learned(d, hard, tr, va, fb, B, Tl)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(fb.cpu().numpy(), redo.astype(np.int32))
    np.testing.assert_array_equal(hard.cpu().numpy(), np.where(redo[:, None], h_p, h_l))
    np.testing.assert_array_equal(tr.cpu().numpy(), np.where(redo, t_p, t_l))
    np.testing.assert_array_equal(va.cpu().numpy(), np.where(redo, v_p, v_l))


@pytest.mark.parametrize("name,fs,rate,B,T,quant", [("s2_c4", 1, "1/2", 3, 3, False),
                                                    ("s2_b6", 0, "2/3", 1, 2, False),
                                                    ("s2_c4", 1, "1/2", 3, 4, True),
                                                    ("s2_b6", 0, "2/3", 2, 3, True)])
def test_training_kernels_vs_torch_reference(name, fs, rate, B, T, quant):
    torch = _torch()
    from tests.learn_ref import TorchLayered
    ctx = ca.FecContext(fs, rate, "QPSK", max_trials=50, max_batch=B)
    q, N = ctx.info["q"], ctx.info["n_ldpc"]
    _, _, cw = ca.gen_frames(fs, rate, "QPSK", 3, B, 1000.0, want_cw=True)
    bits = np.unpackbits(cw, axis=1)[:, :N]
    rng = np.random.default_rng(8)
    # Exact inputs: LLRs and beta in quarters, alpha in {1, 2}.  Every
    # forward value (x, y, APP) is then a multiple of 1/4 below 2^12 in
    # magnitude, which fp32 holds exactly whatever the order of the summed
    # updates, so every min / sign / relu / clamp decision of the kernel is
    # the reference's (ties break towards the first edge in both; d|x|/dx = 0
    # at x = 0 in both).  A scale like 3/4 would add fractional bits with
    # every layer that reads an APP an earlier layer of the iteration wrote.
    llr = np.rint(4 * ((1 - 2 * bits.astype(np.float64)) * 3.0 + rng.normal(0, 3.5, bits.shape))) / 4
    alpha = rng.choice([1.0, 1.0, 2.0], (T, q))
    beta = rng.integers(0, 7, (T, q)) / 4.0
    if quant:
        # the integer decoder's check node: integer LLRs, any alpha / beta
        # (rounded inside), every forward value an exact integer
        llr = np.rint((1 - 2 * bits.astype(np.float64)) * 9.0 + rng.normal(0, 8.0, bits.shape))
        alpha = rng.uniform(0.7, 1.3, (T, q))
        beta = rng.uniform(0.0, 2.6, (T, q))
    kappa = 0.5

    ref = TorchLayered(name)
    a_t = torch.tensor(alpha, requires_grad=True)
    b_t = torch.tensor(beta, requires_grad=True)
    x_t = torch.tensor(llr, requires_grad=True)
    L_r, losses_r, loss_r = ref.run(x_t, torch.from_numpy(bits), a_t, b_t, kappa, quant=quant)
    ga_r, gb_r, gx_r = torch.autograd.grad(loss_r, (a_t, b_t, x_t))

    lc = ca.LearnContext(ctx, B, T)
    f32 = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")
    llr_d, al_d, be_d = f32(llr), f32(alpha), f32(beta)
    cw_d = torch.from_numpy(cw).cuda()
    ga, gb = torch.zeros_like(al_d), torch.zeros_like(be_d)
    loss = torch.zeros(T, dtype=torch.float32, device="cuda")
    app = torch.zeros_like(llr_d)
    gx = torch.zeros_like(llr_d)
    lc.fwd_bwd(llr_d, cw_d, B, al_d, be_d, kappa, ga, gb, loss, app, gx, quantized=quant)
    torch.cuda.synchronize()

    # forward: exact (see above)
    np.testing.assert_array_equal(app.cpu().numpy(), L_r.detach().numpy())
    # losses: fp32 softplus summed over B N bits vs fp64
    np.testing.assert_allclose(loss.cpu().numpy(), losses_r.detach().numpy(), rtol=2e-5)

    def close(got, want, tol):
        got, want = got.cpu().numpy().ravel(), want.numpy().ravel()
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < tol, err
    # gradients: fp32 sigmoid and atomically summed adjoints vs fp64,
    # relative L2 error
    close(ga, ga_r, 1e-4)
    close(gb, gb_r, 1e-4)
    close(gx, gx_r, 1e-4)
    lc.close()
