"""GPU parity: the HIP demap -> LDPC -> BCH path (through the C ABI) against
the CPU oracle on identical seeded inputs.  Integer outputs must match bit
for bit: LLRs, hard bits, remaining-trial counts, BCH flags."""
import numpy as np
import pytest

import cusat_amd as ca
from oracle import oracle_ctypes as oc
from tests.fec_util import (codeword_llr, gpu_bch, gpu_chain, gpu_decode, oracle_code,
                            oracle_demap, snr_scales)

pytestmark = pytest.mark.gpu


def _chain_parity(framesize, rate, mod, es, B, seed, T=50, flags=0):
    ctx = ca.FecContext(framesize, rate, mod, max_trials=T, max_batch=B, flags=flags)
    sym, info, cw = ca.gen_frames(framesize, rate, mod, seed, B, es, want_cw=True)
    nv, gain = snr_scales(es)
    print("case", framesize, rate, mod, es, B, seed, flush=True)
    llr_g, hard_g, tr_g, va_g = gpu_chain(ctx, sym, nv, gain)
    llr_o = oracle_demap(ctx, sym, nv, gain)
    np.testing.assert_array_equal(llr_g, llr_o)
    code = oracle_code(framesize, rate, use_sched=not (flags & ca.CTX_SEQUENTIAL_SCHEDULE))
    hard_o, tr_o = code.decode(llr_o, T)
    np.testing.assert_array_equal(tr_g, tr_o)
    np.testing.assert_array_equal(hard_g, hard_o)
    va_o = oc.bch_check(hard_o, ctx.info["n_bch"], framesize, ctx.info["bch_t"])
    np.testing.assert_array_equal(va_g, va_o)
    return tr_o, va_o, hard_o, info


def test_cfg1_qpsk_1_2_normal():
    """QPSK 1/2 normal, 32 frames at Es/N0 1.5 dB, T = 50."""
    tr, va, _, _ = _chain_parity(0, "1/2", "QPSK", 1.5, 32, 42)
    assert (tr < 0).any() and (tr >= 0).any()  # a mix of failures and successes


@pytest.mark.parametrize("es", [6.5, 7.0, 7.5, 8.0])
def test_cfg2_8psk_3_4_normal(es):
    """8PSK 3/4 normal at the sweep points of the at-scale workload (+ 8.0 dB
    where frames converge)."""
    _chain_parity(0, "3/4", "8PSK", es, 48, 1000 + int(es * 10))


def test_cfg4_16apsk_2_3_normal():
    """16APSK 2/3 (gamma 3.15) normal: log-MAP APSK demap + decode."""
    _chain_parity(0, "2/3", "16APSK", 9.2, 24, 77)


@pytest.mark.parametrize("rate,mod,es", [("5/6", "16APSK", 12.2), ("3/4", "32APSK", 13.4),
                                         ("9/10", "32APSK", 16.5), ("3/5", "8PSK", 6.3),
                                         ("9/10", "8PSK", 11.5), ("1/4", "QPSK", 0.4)])
def test_mixed_modcods_normal(rate, mod, es):
    """The normal-frame MODCODs of the mixed ACM stream, near threshold;
    5/6 and 9/10 carry lane-pair collisions in their schedules."""
    _chain_parity(0, rate, mod, es, 16, 5)


@pytest.mark.parametrize("rate,mod,es", [("1/2", "QPSK", 2.0), ("3/5", "8PSK", 6.5),
                                         ("2/3", "16APSK", 9.8), ("3/4", "32APSK", 13.8),
                                         ("1/4", "QPSK", 0.0), ("8/9", "QPSK", 7.0)])
def test_short_frames(rate, mod, es):
    """Short FECFRAMEs (16200): sequential one-layer schedule, intra-layer
    collisions on every DVB-S2 short code."""
    _chain_parity(1, rate, mod, es, 32, 9)


@pytest.mark.parametrize("fs,rate,mod,es", [(0, "135/180", "128APSK", 18.2), (0, "154/180", "QPSK", 6.4),
                                             (0, "96/180", "8PSK", 5.0), (0, "22/30", "16APSK", 10.6),
                                             (0, "7/9", "32APSK", 12.9), (1, "7/15", "QPSK", 2.0)])
def test_s2x_codes_with_uneven_layers(fs, rate, mod, es):
    """DVB-S2X codes whose layers hold different numbers of data circulants
    (the guarded decoder variants, CMAX 12 / 18 / 30, two-lane and one-lane
    schedules), on the split-edge kernel and on the 6-wave kernel."""
    _chain_parity(fs, rate, mod, es, 12, 21)
    _chain_parity(fs, rate, mod, es, 12, 21, flags=ca.CTX_NO_SPLIT_KERNEL)


@pytest.mark.parametrize("fs,rate,mod,es", [(0, "4/5", "64APSK", 16.4), (0, "135/180", "256APSK", 23.6),
                                             (0, "132/180", "64APSK", 15.3), (0, "100/180", "8APSK", 5.2),
                                             (0, "20/30", "256APSK", 18.4), (0, "22/30", "256APSK", 20.1),
                                             (0, "128/180", "256APSK", 22.6)])
def test_s2x_apsk_rings(fs, rate, mod, es):
    """DVB-S2X constellations of the combined stream and their neighbours,
    at threshold + 0.3 dB (acm_config.h:68,109-121): 64APSK 4/5 (8+16+20+20
    rings, constellation.cc:310-318) and 11/15 = 132/180 (4+12+20+28), 8APSK 5/9,
    256APSK 3/4 and 32/45 (ring form, constellation.cc:400) and the 256APSK-L
    rates 2/3 and 11/15 (explicit Table 15d points); n_mod 6 and 8 with 6 and
    8 interleaver columns.  LLRs of the HIP demapper bit-exact vs the oracle,
    then the decode and BCH flags."""
    tr, va, _, _ = _chain_parity(fs, rate, mod, es, 12, 61)
    assert (tr >= 0).sum() >= 6, "near threshold most frames converge"


@pytest.mark.parametrize("fs,rate,mod,es", [(0, "3/4", "8PSK", 20.0), (0, "2/3", "16APSK", 22.0),
                                             (0, "3/4", "32APSK", 25.0), (1, "5/6", "16APSK", 24.0)])
def test_demap_gain_cap_active(fs, rate, mod, es):
    """High SNR, where the demapper lowers llr_gain to 2 * 320 / dmax
    (capped_llr_gain, demapper_kernel.cu:283-291): 8PSK cap 160, 16APSK 2/3
    cap 349.8, against llr_gain = min(4 / N0, 1000) = 400 .. 1000.  LLRs
    bit-exact vs the oracle with the scaling recomputed from the
    reference's definition, then the decode."""
    from tests.fec_util import reference_demap_scales
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=4)
    nv, gain = snr_scales(es)
    _, _, g = reference_demap_scales(ctx, nv, gain)
    assert g < gain, "the cap is active here"
    tr, va, _, _ = _chain_parity(fs, rate, mod, es, 8, 71)
    assert (tr >= 45).all() and (va == 1).all(), "frames this far above threshold converge at once"


def test_sequential_schedule_flag():
    """CUSAT_CTX_SEQUENTIAL_SCHEDULE = the reference's CUSAT_LDPC_SEQ=1."""
    _chain_parity(0, "3/4", "8PSK", 8.0, 16, 3, flags=ca.CTX_SEQUENTIAL_SCHEDULE)


@pytest.mark.parametrize("rate", ["1/4", "1/3", "2/5", "1/2", "3/5", "2/3", "3/4", "4/5", "5/6",
                                  "8/9", "9/10"])
def test_noiseless_roundtrip_all_normal_rates(rate):
    """Noiseless +-127 codewords converge before the first iteration:
    trials = T and hard bits = the BCH codeword (qa_ldpc_encode.cc:68-86)."""
    T = 50
    ctx = ca.FecContext(0, rate, "QPSK", max_trials=T, max_batch=4)
    _, info, cw = ca.gen_frames(0, rate, "QPSK", 42, 4, 1000.0, want_cw=True)
    llr = codeword_llr(cw, ctx.info["n_ldpc"])
    hard, tr = gpu_decode(ctx, llr)
    assert (tr == T).all()
    np.testing.assert_array_equal(hard, cw[:, :ctx.info["k_ldpc"] // 8])
    np.testing.assert_array_equal(gpu_bch(ctx, hard), np.ones(4, np.int32))


def test_saturation_rails():
    """int16 rails: LLRs at +-32767 / -32768 mixed with noise exercise the
    saturating read, data and parity updates (ldpc_decode_kernel.cu:50-58)."""
    T = 30
    ctx = ca.FecContext(0, "1/2", "QPSK", max_trials=T, max_batch=12)
    code = oracle_code(0, "1/2")
    rng = np.random.default_rng(5)
    _, _, cw = ca.gen_frames(0, "1/2", "QPSK", 8, 12, 1000.0, want_cw=True)
    llr = codeword_llr(cw, 64800, 32767).astype(np.int32)
    llr[0:4] = rng.integers(-32768, 32768, size=(4, 64800))          # random full range
    flip = rng.random((4, 64800)) < 0.08
    llr[4:8] = np.where(flip[:, :], -llr[4:8], llr[4:8])               # railed + errors
    llr[8:10] = rng.choice([-32768, 32767, -1, 0, 1], size=(2, 64800))
    llr[10] = 0                                                        # erasure frame
    llr[11] = np.where(rng.random(64800) < 0.5, -32768, 32767)
    llr = llr.astype(np.int16)
    hard_g, tr_g = gpu_decode(ctx, llr)
    hard_o, tr_o = code.decode(llr, T)
    np.testing.assert_array_equal(tr_g, tr_o)
    np.testing.assert_array_equal(hard_g, hard_o)


def test_zero_iterations_and_batch_limits():
    """T = 0 only runs the syndrome; batch 0 / above max_batch is refused."""
    ctx = ca.FecContext(0, "3/4", "8PSK", max_trials=0, max_batch=2)
    _, _, cw = ca.gen_frames(0, "3/4", "8PSK", 1, 2, 1000.0, want_cw=True)
    llr = codeword_llr(cw, 64800)
    llr[1, 5] = -llr[1, 5]
    hard, tr = gpu_decode(ctx, llr)
    assert list(tr) == [0, -1]
    import torch
    d = torch.zeros((3, 64800), dtype=torch.int16, device="cuda")
    h = torch.zeros((3, 6075), dtype=torch.uint8, device="cuda")
    t = torch.zeros(3, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError):
        ctx.decode(d, h, t, 3)
    with pytest.raises(RuntimeError):
        ctx.decode(d, h, t, 0)


def test_bch_single_bit_errors():
    """BCH check flags a single flipped bit at {0, k-1, k, n-1}
    (apps/bench_bch_gpu.cc:52-104) and passes clean codewords."""
    ctx = ca.FecContext(0, "3/4", "8PSK", max_trials=1, max_batch=8)
    _, _, cw = ca.gen_frames(0, "3/4", "8PSK", 4, 8, 1000.0, want_cw=True)
    nb = ctx.info["n_bch"] // 8
    frames = cw[:, :nb].copy()
    k, n = ctx.info["k_bch"], ctx.info["n_bch"]
    for f, pos in zip(range(4), [0, k - 1, k, n - 1]):
        frames[f, pos // 8] ^= 0x80 >> (pos % 8)
    v = gpu_bch(ctx, frames)
    np.testing.assert_array_equal(v, [0, 0, 0, 0, 1, 1, 1, 1])
    np.testing.assert_array_equal(v, oc.bch_check(frames, n, 0, 12))


def _all_codes():
    """(framesize, rate) of every LDPC code of the tables, with QPSK at
    2.5 dB above the Shannon limit of the rate: the middle rates converge,
    the lowest and highest mostly do not."""
    import math
    out = []
    for fs in (ca.FECFRAME_NORMAL, ca.FECFRAME_SHORT):
        for r in ca.RATES:
            try:
                ca.geometry(fs, r, "QPSK")
            except Exception:
                continue
            a, b = r.split("/")
            eta = 2.0 * int(a) / int(b)
            out.append((fs, r, round(10.0 * math.log10(2.0 ** eta - 1.0) + 2.5, 2)))
    return out


@pytest.mark.parametrize("fs,rate,es", _all_codes())
def test_every_code_against_oracle(fs, rate, es):
    """All 51 LDPC codes (34 normal, 17 short) of the tables: demap, decode
    and BCH bit-exact against the oracle on two frames each."""
    _chain_parity(fs, rate, "QPSK", es, 2, 31)


@pytest.mark.parametrize("offset", [1, 2, 3])
@pytest.mark.parametrize("fs,rate,mod,es", [(0, "3/4", "8PSK", 7.6), (1, "1/2", "QPSK", 2.0)])
def test_unaligned_llr_input(fs, rate, mod, es, offset):
    """LLRs at an address that is not 8-byte aligned (a view `offset` int16
    elements into a buffer) take the decoder's element-wise load path: the
    same hard bits and trials as the aligned input and as the oracle."""
    import torch
    B, T = 6, 50
    ctx = ca.FecContext(fs, rate, mod, max_trials=T, max_batch=B)
    sym, _, _ = ca.gen_frames(fs, rate, mod, 91 + offset, B, es)
    nv, gain = snr_scales(es)
    llr = oracle_demap(ctx, sym, nv, gain)
    N = llr.shape[1]
    print("case", fs, rate, mod, es, "offset", offset, flush=True)
    buf = torch.zeros(B * N + 8, dtype=torch.int16, device="cuda")
    buf[offset:offset + B * N] = torch.from_numpy(llr.reshape(-1)).cuda()
    view = buf[offset:offset + B * N]
    assert view.data_ptr() % 8 != 0
    hard = torch.zeros((B, ctx.info["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
    tr = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    ctx.decode(view, hard, tr, B)
    torch.cuda.synchronize()
    hard_a, tr_a = gpu_decode(ctx, llr)
    hard_o, tr_o = oracle_code(fs, rate).decode(llr, T)
    np.testing.assert_array_equal(tr_a, tr_o)
    np.testing.assert_array_equal(hard_a, hard_o)
    np.testing.assert_array_equal(tr.cpu().numpy(), tr_o)
    np.testing.assert_array_equal(hard.cpu().numpy(), hard_o)


def test_demap_bench_runs():
    """The per-code APSK demap benchmark runs on every constellation it
    lists (two frames, one timed launch each)."""
    from cusat_amd import demap_bench
    res = demap_bench.run(frames=2, reps=1, log=lambda s: print(s, flush=True))
    assert [r["modcod"][2] for r in res] == [m for _, m, _ in demap_bench.APSK]
    assert all(r["us_per_launch"] > 0 for r in res)
