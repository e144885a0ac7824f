"""Post-BCH baseband step on the GPU (cusat_bb_extract) against the
baseband reference of tests/bb_ref.py: header edge cases of
qa_bbframe_parser.cc (valid GSE, every single-bit header error, DFL not a
byte multiple, DFL too long, TS/GS not generic packetized, DFL 0, largest
DFL), FEC-failed frames, and the whole receive path from GSE payloads."""
import numpy as np
import pytest

import cusat_amd as ca
from tests import bb_ref
from tests.fec_util import snr_scales

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("fs,rate", [(0, "1/2"), (1, "1/2"), (0, "9/10")])
def test_bb_extract_header_cases(fs, rate):
    torch = _torch()
    rng = np.random.default_rng(5)
    g = ca.geometry(fs, rate, "QPSK")
    kbch, kb = g["k_bch"] // 8, g["k_ldpc"] // 8
    prbs = ca.bb_prbs(kbch)
    cases = []
    for n in (min(1000, kbch - 10), 0, 1, kbch - 10, 37):
        cases.append(dict(payload=rng.integers(0, 256, n, dtype=np.uint8)))
    for bit in range(0, 80, 7):
        cases.append(dict(payload=rng.integers(0, 256, 100, dtype=np.uint8), corrupt_bit=bit))
    cases.append(dict(payload=rng.integers(0, 256, 100, dtype=np.uint8), dfl_bits=803))
    cases.append(dict(payload=rng.integers(0, 256, 10, dtype=np.uint8), dfl_bits=8 * (kbch - 10) + 8))
    for ts in (1, 2, 3):
        cases.append(dict(payload=rng.integers(0, 256, 50, dtype=np.uint8), ts_gs=ts))
    B = len(cases)
    hard = np.zeros((B, kb), np.uint8)
    clear = []
    for b, c in enumerate(cases):
        f = bb_ref.make_bbframe(kbch, c["payload"], ts_gs=c.get("ts_gs", 0), dfl_bits=c.get("dfl_bits"),
                                rng=rng, corrupt_bit=c.get("corrupt_bit"))
        clear.append(f)
        hard[b, :kbch] = f ^ prbs
        hard[b, kbch:] = rng.integers(0, 256, kb - kbch, dtype=np.uint8)
    trials = rng.integers(0, 30, B).astype(np.int32)
    valid = np.ones(B, np.int32)
    trials[2] = -1          # LDPC failure
    valid[3] = 0            # BCH failure
    ctx = ca.FecContext(fs, rate, "QPSK", max_trials=10, max_batch=B)
    d = lambda a: torch.from_numpy(a).cuda()
    out = torch.zeros((B, kbch - 10), dtype=torch.uint8, device="cuda")
    dfl = torch.full((B,), -7, dtype=torch.int32, device="cuda")
    ctx.bb_extract(d(hard), d(trials), d(valid), out, dfl, B)
    torch.cuda.synchronize()
    out, dfl = out.cpu().numpy(), dfl.cpu().numpy()
    for b in range(B):
        if b in (2, 3):
            assert dfl[b] == -1
            continue
        want = bb_ref.extract(clear[b], kbch)
        n = 0 if want is None else len(want)
        assert dfl[b] == n, (b, cases[b].get("corrupt_bit"), dfl[b], n)
        assert bytes(out[b, :n]) == (want or b"")


def test_receive_path_recovers_gse_payloads():
    """GSE payloads -> BBHEADER -> BB scramble -> BCH -> LDPC -> 8PSK ->
    AWGN -> demap -> decode -> BCH check -> BB step: payloads back."""
    torch = _torch()
    fs, rate, mod, B, es = 0, "3/4", "8PSK", 6, 9.5
    g = ca.geometry(fs, rate, mod)
    kbch = g["k_bch"] // 8
    prbs = ca.bb_prbs(kbch)
    rng = np.random.default_rng(9)
    pay = [rng.integers(0, 256, n, dtype=np.uint8) for n in (5000, 1, 0, kbch - 10, 321, 4096)]
    info = np.stack([bb_ref.make_bbframe(kbch, p, rng=rng) ^ prbs for p in pay])
    sym, _ = ca.encode_frames(fs, rate, mod, 3, info, es)
    ctx = ca.FecContext(fs, rate, mod, max_trials=50, max_batch=B)
    nv, gain = snr_scales(es)
    llr = torch.empty((B, g["n_ldpc"]), dtype=torch.int16, device="cuda")
    hard = torch.empty((B, g["k_ldpc"] // 8), dtype=torch.uint8, device="cuda")
    tr = torch.empty(B, dtype=torch.int32, device="cuda")
    va = torch.empty(B, dtype=torch.int32, device="cuda")
    out = torch.zeros((B, kbch - 10), dtype=torch.uint8, device="cuda")
    dfl = torch.empty(B, dtype=torch.int32, device="cuda")
    ctx.rx_fec_batch(torch.from_numpy(sym).cuda(), llr, B, nv, gain, hard, tr, va)
    ctx.bb_extract(hard, tr, va, out, dfl, B)
    torch.cuda.synchronize()
    assert (tr.cpu().numpy() >= 0).all() and (va.cpu().numpy() == 1).all()
    dfl, out = dfl.cpu().numpy(), out.cpu().numpy()
    for b, p in enumerate(pay):
        assert dfl[b] == len(p)
        assert bytes(out[b, :len(p)]) == bytes(p)
