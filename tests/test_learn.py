"""Learned offset/normalised min-sum, CPU side.

* The fp32 torch reference of the trainable decoder (tests/learn_ref.py) at
  alpha = 1, beta = 1 on integer LLRs is the plain decoder without rounding:
  on frames that never converge its hard decision after T iterations must
  equal the integer oracle's, on a sequential short code and on a normal code
  with the two-layer schedule (colliding updates included).
* Its autograd gradients agree with central differences.
* The oracle's parameter hooks at identity values reproduce the plain decoder.
* Parameter quantisation of the training module.
"""
import numpy as np
import pytest
import torch

import cusat_amd as ca
from cusat_amd import learn
from oracle import oracle_ctypes as oc
from tests.learn_ref import TorchLayered


def _noisy_llr(fs, rate, B, seed, mean, sigma, integer=True):
    _, _, cw = ca.gen_frames(fs, rate, "QPSK", seed, B, 1000.0, want_cw=True)
    n = 64800 if fs == 0 else 16200
    bits = np.unpackbits(cw, axis=1)[:, :n]
    rng = np.random.default_rng(seed)
    llr = (1 - 2 * bits.astype(np.float64)) * mean + rng.normal(0, sigma, bits.shape)
    if integer:
        llr = np.clip(np.rint(llr), -32767, 32767).astype(np.int16)
    return llr, bits, cw


def _pack(hard_bits):
    return np.packbits(hard_bits.astype(np.uint8), axis=1)


@pytest.mark.parametrize("name,fs,rate,B,T", [("s2_c4", 1, "1/2", 4, 3), ("s2_b6", 0, "2/3", 2, 2)])
def test_torch_reference_is_the_plain_decoder(name, fs, rate, B, T):
    llr, bits, _ = _noisy_llr(fs, rate, B, 7, 3.0, 4.0)
    code = oc.Code(name)
    h_o, r_o = code.decode(llr, T)
    assert np.all(r_o == -1), "frames converged; the comparison needs T full iterations"
    ref = TorchLayered(name)
    q = ref.q
    one = torch.ones((T, q), dtype=torch.float64)
    L, _, _ = ref.run(torch.from_numpy(llr.astype(np.float64)), torch.from_numpy(bits), one, one, 0.5)
    hard = _pack((L[:, : ref.K] < 0).numpy())
    np.testing.assert_array_equal(hard, h_o)
    # a dropped collision term or a wrong parity index changes the result
    L2, _, _ = ref.run(torch.from_numpy(llr.astype(np.float64)), torch.from_numpy(bits), one,
                       one * 2.0, 0.5)
    assert not np.array_equal(_pack((L2[:, : ref.K] < 0).numpy()), h_o)


@pytest.mark.parametrize("name,fs,rate,B,T", [("s2_c4", 1, "1/2", 4, 4), ("s2_b6", 0, "2/3", 2, 3)])
def test_quantised_torch_reference_is_the_integer_decoder(name, fs, rate, B, T):
    """The quantised model with real-valued (alpha, beta) is the integer
    decoder with the rounded parameters of learn.quantize."""
    llr, bits, _ = _noisy_llr(fs, rate, B, 21, 3.0, 4.0)
    ref = TorchLayered(name)
    rng = np.random.default_rng(2)
    alpha = rng.uniform(0.6, 1.4, (T, ref.q))
    beta = rng.uniform(0.0, 2.7, (T, ref.q))
    L, _, _ = ref.run(torch.from_numpy(llr.astype(np.float64)), torch.from_numpy(bits),
                      torch.from_numpy(alpha), torch.from_numpy(beta), 0.5, quant=True)
    bq, aq = learn.quantize(alpha, beta)
    code = oc.Code(name)
    h_o, r_o = code.decode(llr, T, beta=bq, alpha=aq, n_param_iters=T)
    assert np.all(r_o == -1)
    np.testing.assert_array_equal(_pack((L[:, : ref.K] < 0).numpy()), h_o)


def test_torch_reference_gradients_vs_central_differences():
    llr, bits, _ = _noisy_llr(1, "1/2", 1, 3, 2.5, 3.0, integer=False)
    ref = TorchLayered("s2_c4")
    T, q = 2, ref.q
    rng = np.random.default_rng(5)
    alpha = torch.tensor(rng.uniform(0.7, 1.0, (T, q)), dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(rng.uniform(0.0, 1.0, (T, q)), dtype=torch.float64, requires_grad=True)
    x = torch.tensor(llr, dtype=torch.float64, requires_grad=True)
    b = torch.from_numpy(bits)
    _, _, loss = ref.run(x, b, alpha, beta, 0.5)
    ga, gb, gx = torch.autograd.grad(loss, (alpha, beta, x))
    h = 1e-6
    for which in range(3):
        d = [torch.zeros_like(alpha), torch.zeros_like(beta), torch.zeros_like(x)]
        d[which] = torch.tensor(rng.normal(size=d[which].shape), dtype=torch.float64)
        with torch.no_grad():
            fp = ref.run(x + h * d[2], b, alpha + h * d[0], beta + h * d[1], 0.5)[2]
            fm = ref.run(x - h * d[2], b, alpha - h * d[0], beta - h * d[1], 0.5)[2]
        fd = float((fp - fm) / (2 * h))
        an = float((ga * d[0]).sum() + (gb * d[1]).sum() + (gx * d[2]).sum())
        assert abs(fd - an) <= 1e-4 * max(abs(an), 1e-9), (which, fd, an)


@pytest.mark.parametrize("alpha64", [0, 64])
def test_oracle_identity_parameters_are_plain(alpha64):
    llr, _, _ = _noisy_llr(0, "2/3", 3, 11, 4.0, 4.5)
    code = oc.Code("s2_b6")
    h0, r0 = code.decode(llr, 12)
    T = 12
    beta = np.ones((T, code.q), np.int16)
    alpha = np.full((T, code.q), alpha64, np.int16)
    h1, r1 = code.decode(llr, 12, beta=beta, alpha=alpha, n_param_iters=T)
    np.testing.assert_array_equal(h0, h1)
    np.testing.assert_array_equal(r0, r1)
    # and a real offset changes the trajectory
    h2, r2 = code.decode(llr, 12, beta=beta * 3, alpha=alpha, n_param_iters=T)
    assert not (np.array_equal(h0, h2) and np.array_equal(r0, r2))


def test_quantisation():
    alpha = np.array([[0.5, 1.0, 1.26, 20.0]], np.float32)
    beta = np.array([[-1.0, 0.49, 1.5, 400.0]], np.float32)
    b, a = learn.quantize(alpha, beta)
    assert b.dtype == np.int16 and a.dtype == np.int16
    assert list(a[0]) == [32, 64, 81, 1023]
    assert list(b[0]) == [0, 0, 2, 255]


def test_trained_parameter_files():
    """Every shipped parameter file fits its code (n_iters x q int16 tables
    in the decoder hooks' ranges), the lookup finds exactly the files whose
    own evaluation found a gain, and no two files claim one code."""
    import glob
    import json
    import os
    seen = set()
    for f in sorted(glob.glob(os.path.join(learn.LEARNED_DIR, "*.json"))):
        d = json.load(open(f))
        code = (d["framesize"], d["rate"], d["modulation"])
        assert code not in seen, f
        seen.add(code)
        q = ca.geometry(*code)["q"]
        b, a = np.array(d["beta_q"]), np.array(d["alpha_q"])
        assert b.shape == a.shape == (d["n_iters"], q), f
        assert b.min() >= 0 and b.max() <= 255 and a.min() >= 1 and a.max() <= 1023, f
        # the stored fp32 values are rounded to 5 decimals: a value on a
        # rounding boundary may quantise one step away
        bq, aq = learn.quantize(d["alpha"], d["beta"])
        assert np.abs(bq.astype(int) - b).max() <= 1 and np.mean(bq != b) < 0.01, f
        assert np.abs(aq.astype(int) - a).max() <= 1 and np.mean(aq != a) < 0.01, f
        found = learn.params_for(*code)
        assert (found == f) == d.get("use", True), f
    # every code of the s2x stream has been trained
    want = {(fs, r, m) for r, m, _ in learn.MODCODS12 + learn.S2X
            for fs in ((0, 1) if learn.has_short(r, m) else (0,))}
    assert want <= seen


def test_packed_alpha_scale_identity():
    """The decoders evaluate min((m n + 32) >> 6, 127) in 16-bit lanes after
    clamping m to ceil(8096 / n) (alpha word n | bound << 16, packed by
    cusat_ldpc_set_params): exhaustively equal to the 32-bit form for every
    scale n in 1..1023 and magnitude m in 0..32767."""
    m = np.arange(32768, dtype=np.int64)
    for n in range(1, 1024):
        ref = np.minimum((m * n + 32) >> 6, 127)
        bound = (8096 + n - 1) // n
        mc = np.minimum(m, bound)
        assert (mc * n + 32).max() < 1 << 16
        np.testing.assert_array_equal(np.minimum(((mc * n + 32) & 0xffff) >> 6, 127), ref)
