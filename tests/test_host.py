"""Host-side tests (no GPU): the C ABI library loads and exports every
declared symbol, the host tables agree value for value with the reference's
own host code (oracle/_ref, when built), and the seeded generator produces
valid, reproducible codewords."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import cusat_amd as ca
from oracle import oracle_ctypes as oc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_host.so")


def declared_symbols():
    syms = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            syms += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(cusat_\w+)\s*\(", text, re.M)
    return sorted(set(syms))


def test_library_exports_every_declared_symbol():
    L = ca.lib()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(L, s), s


def test_errors_are_reported_not_fatal():
    inf = ca.FecInfo()
    rc = ca.lib().cusat_fec_query(1, ca.rate_id("9/10"), 0, 0, ctypes.byref(inf))
    assert rc == -1
    assert b"unsupported" in ca.lib().cusat_last_error()
    with pytest.raises(RuntimeError):
        ca.geometry(0, "3/4", 9)  # no such modulation


SUPPORTED = [(0, r) for r in ["1/4", "1/3", "2/5", "1/2", "3/5", "2/3", "3/4", "4/5", "5/6", "8/9",
                              "9/10", "13/45", "9/20", "11/20", "26/45", "28/45", "23/36", "25/36",
                              "13/18", "7/9", "90/180", "96/180", "100/180", "104/180", "116/180",
                              "124/180", "128/180", "132/180", "135/180", "140/180", "154/180",
                              "18/30", "20/30", "22/30"]] + \
            [(1, r) for r in ["1/4", "1/3", "2/5", "1/2", "3/5", "2/3", "3/4", "4/5", "5/6", "8/9",
                              "11/45", "4/15", "14/45", "7/15", "8/15", "26/45", "32/45"]]


@pytest.mark.parametrize("fs,rate", SUPPORTED)
def test_query_every_code(fs, rate):
    g = ca.geometry(fs, rate, "QPSK")
    n = 64800 if fs == 0 else 16200
    assert g["n_ldpc"] == n and g["n_bch"] == g["k_ldpc"]
    assert g["k_bch"] + g["bch_degree"] == g["n_bch"]
    assert g["lds_bytes"] <= 160 * 1024
    assert g["q"] * 360 == n - g["k_ldpc"]


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L = ctypes.CDLL(REF_LIB)
    i, vp = ctypes.c_int, ctypes.c_void_p
    L.ref_constellation.argtypes = [i, i, i, vp, ctypes.POINTER(i)]
    L.ref_interleaver.argtypes = [i, i, i, vp, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.ref_fec_params.argtypes = [i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
    L.ref_bch.argtypes = [i, i, i, vp, vp, vp]
    L.ref_ldpc_table.argtypes = [i, i] + [ctypes.POINTER(i)] * 5 + [vp, i]
    return L


@pytest.mark.parametrize("fs,rate", SUPPORTED)
def test_fec_params_and_table_selection_vs_reference(fs, rate):
    L = _ref()
    k, n, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.ref_fec_params(fs, ca.rate_id(rate), ctypes.byref(k), ctypes.byref(n), ctypes.byref(t)) == 0
    g = ca.geometry(fs, rate, "QPSK")
    assert (g["k_bch"], g["n_bch"], g["bch_t"]) == (k.value, n.value, t.value)
    v = [ctypes.c_int() for _ in range(5)]
    addr = np.zeros(20000, np.uint16)
    total = L.ref_ldpc_table(fs, ca.rate_id(rate), *[ctypes.byref(x) for x in v],
                             addr.ctypes.data_as(ctypes.c_void_p), len(addr))
    N, K, n_rows, lt, lmc = [x.value for x in v]
    assert (N, K) == (g["n_ldpc"], g["k_ldpc"])
    assert g["lt_alloc"] == lt + 1 and g["cnl"] == lmc - 2
    # the same addresses as our embedded table text, in order
    from tests.test_tables import TABLES
    mine = [t for t in TABLES.values() if t["N"] == N and t["K"] == K and
            [a for r in t["rows"] for a in r] == list(addr[:total])]
    assert len(mine) >= 1


MODCODS = [("QPSK", "1/2"), ("8PSK", "3/4"), ("8PSK", "3/5"), ("8PSK", "9/10"), ("16APSK", "2/3"),
           ("16APSK", "5/6"), ("32APSK", "3/4"), ("32APSK", "9/10"), ("64APSK", "4/5"),
           ("128APSK", "135/180"), ("256APSK", "135/180"), ("8APSK", "100/180"),
           ("16APSK", "90/180"), ("16APSK", "18/30"), ("32APSK", "2/3"), ("32APSK", "128/180"),
           ("64APSK", "132/180"), ("64APSK", "128/180"), ("128APSK", "140/180"),
           ("256APSK", "116/180"), ("256APSK", "20/30"), ("256APSK", "22/30")]


@pytest.mark.parametrize("mod,rate", MODCODS)
@pytest.mark.parametrize("fs", [0, 1])
def test_constellation_and_interleaver_vs_reference(mod, rate, fs):
    if fs == 1 and (rate in ("9/10", "135/180") or mod in ("64APSK", "128APSK", "256APSK")):
        pytest.skip("no short frame for this MODCOD")
    L = _ref()
    pts, col, nm, rows = ca.modcod_constellation(fs, rate, mod)
    rp = np.zeros(512, np.float32)
    bits = ctypes.c_int()
    assert L.ref_constellation(ca.mod_id(mod), ca.rate_id(rate), fs,
                               rp.ctypes.data_as(ctypes.c_void_p), ctypes.byref(bits)) == 0
    assert bits.value == nm
    np.testing.assert_array_equal(pts.view(np.uint32), rp[: len(pts)].view(np.uint32))
    rc = np.zeros(8, np.uint32)
    nc, rr = ctypes.c_int(), ctypes.c_int()
    assert L.ref_interleaver(ca.mod_id(mod), ca.rate_id(rate), fs, rc.ctypes.data_as(ctypes.c_void_p),
                             ctypes.byref(nc), ctypes.byref(rr)) == 0
    np.testing.assert_array_equal(col, rc)
    assert rows == rr.value


@pytest.mark.parametrize("fs,rate", [(0, "1/2"), (0, "2/3"), (0, "3/4"), (0, "9/10"), (1, "1/2"),
                                     (1, "8/9")])
def test_bch_encoder_vs_reference(fs, rate):
    L = _ref()
    g = ca.geometry(fs, rate, "QPSK")
    rng = np.random.default_rng(11)
    msg = rng.integers(0, 256, (2, g["k_bch"] // 8), dtype=np.uint8)
    mine = ca.bch_encode(fs, rate, msg)
    gen = np.zeros(4, np.uint64)
    for b in range(2):
        cw = np.zeros(g["n_bch"] // 8, np.uint8)
        p = L.ref_bch(fs, g["n_bch"], g["bch_t"], msg[b].ctypes.data_as(ctypes.c_void_p),
                      cw.ctypes.data_as(ctypes.c_void_p), gen.ctypes.data_as(ctypes.c_void_p))
        assert p == g["bch_degree"]
        np.testing.assert_array_equal(mine[b], cw)
    d, go = oc.bch_generator(fs, g["bch_t"])
    np.testing.assert_array_equal(go, gen)


@pytest.mark.parametrize("mod,rate,fs,order,rows", [tuple(v) for v in KAT["interleaver"]])
def test_interleaver_kats(mod, rate, fs, order, rows):
    _, col, nm, r = ca.modcod_constellation(fs, rate, mod)
    if rows is not None:
        assert r == rows
    if order is not None:
        assert [int(c) // r for c in col[:nm]] == order


@pytest.mark.parametrize("mod,rate", MODCODS)
def test_constellation_properties(mod, rate):
    pts, _, nm, _ = ca.modcod_constellation(0, rate, mod)
    P = pts.reshape(-1, 2).astype(np.float64)
    assert len(P) == 1 << nm
    assert abs((P ** 2).sum(1).mean() - 1.0) < 1e-3           # unit average energy
    assert len({(round(x, 5), round(y, 5)) for x, y in P}) == len(P)  # distinct
    if mod == "QPSK":
        a = np.float32(np.sqrt(0.5))
        np.testing.assert_array_equal(pts, np.float32([a, a, a, -a, -a, a, -a, -a]))


@pytest.mark.parametrize("rate,g1,g2", [tuple(v) for v in KAT["apsk32_ratios"]])
def test_32apsk_ring_ratios(rate, g1, g2):
    """qa_constellation.cc:102-128: labels 17 / 0 / 8 sit on rings 1 / 2 / 3."""
    pts, _, _, _ = ca.modcod_constellation(0, rate, "32APSK")
    r = np.hypot(pts[0::2].astype(np.float64), pts[1::2].astype(np.float64))
    assert abs(r[0] / r[17] - g1) < 1e-6 * g1 and abs(r[8] / r[17] - g2) < 1e-6 * g2
    assert (np.abs(r - r[17]) < 1e-5).sum() == 4 and (np.abs(r - r[0]) < 1e-5).sum() == 12


def test_generator_reproducible_and_valid():
    s1, i1, c1 = ca.gen_frames(0, "3/4", "8PSK", 5, 6, 7.0, want_cw=True)
    s2, i2, c2 = ca.gen_frames(0, "3/4", "8PSK", 5, 6, 7.0, want_cw=True)
    s3, _, _ = ca.gen_frames(0, "3/4", "8PSK", 6, 6, 7.0)
    assert np.array_equal(s1, s2) and np.array_equal(i1, i2) and not np.array_equal(s1, s3)
    # frame b does not depend on the batch size
    s4, i4, _ = ca.gen_frames(0, "3/4", "8PSK", 5, 2, 7.0)
    assert np.array_equal(s4, s1[:2]) and np.array_equal(i4, i1[:2])
    code = oc.Code("s2_b7")
    for b in range(6):
        bits = np.unpackbits(c1[b])
        llr = np.where(bits == 0, 100, -100).astype(np.int16)
        assert code.syndrome_ok(llr)
        assert np.array_equal(c1[b, : 48408 // 8], i1[b])
    assert list(oc.bch_check(c1[:, :6075], 48600, 0, 12)) == [1] * 6
    # noise level: Es/N0 7 dB on unit-energy 8PSK
    n0 = 10 ** (-0.7)
    sn, _, _ = ca.gen_frames(0, "3/4", "8PSK", 5, 6, 1000.0)
    var = ((s1 - sn) ** 2).sum(2).mean()
    assert abs(var / n0 - 1) < 0.02


def test_generator_128apsk_padding_and_stuffing():
    sym, _, cw = ca.gen_frames(0, "135/180", "128APSK", 1, 1, 1000.0, want_cw=True)
    g = ca.geometry(0, "135/180", "128APSK")
    assert sym.shape[1] == 9270 and g["demap_len"] == 9258
    pts, _, _, _ = ca.modcod_constellation(0, "135/180", "128APSK")
    assert np.all(sym[0, 9258:] == pts.reshape(-1, 2)[127])


def test_frame_acceptance_truth_table():
    """lib/qa_rx_fec_chain.cc:9-17: accept only a converged LDPC with a valid BCH."""
    t = np.array([25, 0, -1, 3, -1])
    v = np.array([1, 1, 1, 0, 0])
    assert list(ca.frame_passes_fec_checks(t, v)) == [True, True, False, False, False]


@pytest.mark.parametrize("workload,n,short,mods", [("mixed12", 2048, 0.30, 12), ("s2x", 4096, 0.25, 15)])
def test_stream_workload_spec(workload, n, short, mods):
    from cusat_amd.stream_bench import make_stream
    spec = make_stream(workload, n, 1)
    assert len(spec) == n and len({(r, m) for _, r, m, _ in spec}) == mods
    frac = np.mean([fs == 1 for fs, _, _, _ in spec])
    assert abs(frac - short) < 0.03
    # short frames only where a short code exists
    for fs, r, m, _ in spec:
        if fs == 1:
            ca.geometry(1, r, m)
            assert r not in ("9/10", "135/180")


def test_bb_prbs_matches_register_definition():
    """qa_bb_scrambler.cc:38-57: first output bits 0 0 0 0 0 0 1 1; every
    byte equals the bit-serial register of figure 5, MSB first."""
    from tests.bb_ref import prbs_bits
    n = 7274
    seq = ca.bb_prbs(n)
    assert seq[0] == 0x03
    bits = prbs_bits(8 * n)
    assert list(bits[:8]) == [0, 0, 0, 0, 0, 0, 1, 1]
    np.testing.assert_array_equal(seq, np.packbits(bits))


def test_bb_crc8_kats():
    """qa_crc8.cc:12-29: CRC-8/DVB-S2 check value, appended CRC checks to 0,
    single-bit errors detected."""
    from tests.bb_ref import crc8
    assert ca.bb_crc8(np.frombuffer(b"123456789", np.uint8)) == 0xBC
    hdr = np.array([0xF2, 0x00, 0x00, 0x00, 0x7E, 0x10, 0x00, 0x00, 0x00], np.uint8)
    full = np.append(hdr, ca.bb_crc8(hdr)).astype(np.uint8)
    assert ca.bb_crc8(full) == 0
    for bit in range(80):
        f = full.copy()
        f[bit // 8] ^= 0x80 >> (bit % 8)
        assert ca.bb_crc8(f) != 0
    rng = np.random.default_rng(4)
    for _ in range(50):
        d = rng.integers(0, 256, int(rng.integers(0, 40)), dtype=np.uint8)
        assert ca.bb_crc8(d) == crc8(d)


def test_encode_frames_is_gen_frames_chain():
    """cusat_encode_frames on the info bytes gen_frames drew gives the same
    codewords (same BCH / LDPC / map), noiseless."""
    s1, info, cw1 = ca.gen_frames(0, "2/3", "16APSK", 9, 3, 1000.0, want_cw=True)
    s2, cw2 = ca.encode_frames(0, "2/3", "16APSK", 9, info, 1000.0, want_cw=True)
    np.testing.assert_array_equal(cw1, cw2)
    np.testing.assert_array_equal(s1, s2)


def test_dd_snr_restatement_equals_the_library_host_estimate():
    """The float32 restatement of the decision-directed SNR the GPU tests
    check against (tests/fec_util.dd_snr_host) equals libcusat_dsp's host
    estimate (qpsk_snr_dd / constellation_snr_dd) bit for bit."""
    from tests.fec_util import dd_snr_host
    L = ctypes.CDLL(os.path.join(os.path.dirname(ca.__file__), "libcusat_dsp.so"))
    q = L._ZN9cusat_dsp11qpsk_snr_ddEPKSt7complexIfEi
    q.restype, q.argtypes = ctypes.c_float, [ctypes.c_void_p, ctypes.c_int]
    c = L._ZN9cusat_dsp20constellation_snr_ddEPKSt7complexIfEiS3_ii
    c.restype = ctypes.c_float
    c.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    rng = np.random.default_rng(4)
    pts16 = ca.modcod_constellation(0, "2/3", "16APSK")[0].reshape(-1, 2).astype(np.float32)
    for sigma in (0.05, 0.4, 2.0):
        sym = (rng.choice([-0.70710677, 0.70710677], (3000, 2)) + rng.normal(0, sigma, (3000, 2))).astype(np.float32)
        got = q(sym.ctypes.data, len(sym))
        assert np.float32(got).view(np.uint32) == dd_snr_host(sym, None, True).view(np.uint32)
        sym = (pts16[rng.integers(0, 16, 4001)] + rng.normal(0, sigma, (4001, 2))).astype(np.float32)
        got = c(pts16.ctypes.data, 16, sym.ctypes.data, len(sym), 8)
        assert np.float32(got).view(np.uint32) == dd_snr_host(sym, pts16, False).view(np.uint32)
    zero = np.zeros((64, 2), np.float32)
    assert q(zero.ctypes.data, 64) == dd_snr_host(zero, None, True)
    bad = zero.copy()
    bad[8, 0] = np.nan
    assert c(pts16.ctypes.data, 16, bad.ctypes.data, 64, 8) == dd_snr_host(bad, pts16, False) == 0.0


def _ref_ext():
    L = _ref()
    i, vp, f = ctypes.c_int, ctypes.c_void_p, ctypes.c_float
    L.ref_qpsk_snr_dd.restype = f
    L.ref_qpsk_snr_dd.argtypes = [vp, i]
    L.ref_constellation_snr_dd.restype = f
    L.ref_constellation_snr_dd.argtypes = [i, i, i, vp, i, i]
    L.ref_bb_scramble.argtypes = [vp, vp, i]
    L.ref_crc8.argtypes = [vp, i]
    L.ref_extract_gse_datafield.argtypes = [vp, i, vp, i]
    L.ref_build_gse_bbframe.argtypes = [i, i, vp, i, vp]
    return L


def _dsp_snr():
    L = ctypes.CDLL(os.path.join(os.path.dirname(ca.__file__), "libcusat_dsp.so"))
    q = L._ZN9cusat_dsp11qpsk_snr_ddEPKSt7complexIfEi
    q.restype, q.argtypes = ctypes.c_float, [ctypes.c_void_p, ctypes.c_int]
    c = L._ZN9cusat_dsp20constellation_snr_ddEPKSt7complexIfEiS3_ii
    c.restype = ctypes.c_float
    c.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    return q, c


@pytest.mark.parametrize("mod,rate", [("QPSK", "1/2"), ("8PSK", "3/4"), ("16APSK", "2/3"), ("32APSK", "3/4"),
                                      ("64APSK", "4/5"), ("128APSK", "135/180"), ("256APSK", "135/180"),
                                      ("256APSK", "20/30"), ("8APSK", "100/180")])
def test_dd_snr_host_estimate_vs_reference(mod, rate):
    """libcusat_dsp's decision-directed SNR estimates (the chain's DD re-
    estimate) equal the reference's own modulation/dd_snr.cc, compiled from
    its source, bit for bit: QPSK over every symbol with sign decisions,
    other constellations over every 8th symbol with the nearest point; noisy
    frames at three noise levels, all-zero, NaN, inf and huge symbols."""
    R = _ref_ext()
    q, c = _dsp_snr()
    pts = ca.modcod_constellation(0, rate, mod)[0].reshape(-1, 2).astype(np.float32)
    mid, rid = ca.mod_id(mod), ca.rate_id(rate)
    rng = np.random.default_rng(12)
    n = 4001
    cases = []
    for sigma in (0.03, 0.2, 1.5):
        cases.append((pts[rng.integers(0, len(pts), n)] + rng.normal(0, sigma, (n, 2))).astype(np.float32))
    cases.append(np.zeros((n, 2), np.float32))
    bad = cases[0].copy()
    bad[16, 1] = np.nan
    cases.append(bad)
    big = cases[1].copy()
    big[8, 0] = np.inf
    cases.append(big)
    cases.append((cases[1] * np.float32(1e19)).astype(np.float32))
    for sym in cases:
        sym = np.ascontiguousarray(sym)
        if mod == "QPSK":
            got, want = q(sym.ctypes.data, n), R.ref_qpsk_snr_dd(sym.ctypes.data, n)
        else:
            got = c(pts.ctypes.data, len(pts), sym.ctypes.data, n, 8)
            want = R.ref_constellation_snr_dd(mid, rid, 0, sym.ctypes.data, n, 8)
        assert np.float32(got).view(np.uint32) == np.float32(want).view(np.uint32), (got, want)


@pytest.mark.parametrize("mod,rate", MODCODS)
def test_demapper_dmax_and_gain_cap_vs_reference(mod, rate):
    """The demapper's loudest-bit distance and capped gain
    (capped_llr_gain, demapper_kernel.cu:216-239,283-291) recomputed from the
    reference's own constellation (oracle/_ref) by the reference's
    definition equal what the library uses (cusat_fec_query)."""
    L = _ref()
    rp = np.zeros(512, np.float32)
    bits = ctypes.c_int()
    assert L.ref_constellation(ca.mod_id(mod), ca.rate_id(rate), 0, rp.ctypes.data_as(ctypes.c_void_p),
                               ctypes.byref(bits)) == 0
    P = rp[: 2 << bits.value].reshape(-1, 2)
    from tests.fec_util import dmax_by_definition
    want = dmax_by_definition(P, bits.value, mod)
    g = ca.geometry(0, rate, mod)
    assert np.float32(g["llr_dmax"]).view(np.uint32) == want.view(np.uint32), (g["llr_dmax"], want)
    cap = np.float32(np.float32(2.0) * np.float32(320.0)) / want
    assert np.float32(g["llr_gain_cap"]).view(np.uint32) == np.float32(cap).view(np.uint32)
    if (mod, rate) == ("16APSK", "2/3"):
        assert abs(float(want) - 1.8295) < 1e-3 and abs(float(cap) - 349.8) < 0.1


def test_bb_scrambler_vs_reference():
    """The PRBS the library descrambles with equals the reference's
    bb_scrambler.cc over a whole BBFRAME (7274 bytes), and scrambling is XOR."""
    R = _ref_ext()
    n = 7274
    zero = np.zeros(n, np.uint8)
    seq = np.zeros(n, np.uint8)
    assert R.ref_bb_scramble(zero.ctypes.data, seq.ctypes.data, n) == 0
    np.testing.assert_array_equal(ca.bb_prbs(n), seq)
    data = np.random.default_rng(3).integers(0, 256, n, dtype=np.uint8)
    out = np.zeros(n, np.uint8)
    R.ref_bb_scramble(data.ctypes.data, out.ctypes.data, n)
    np.testing.assert_array_equal(out, data ^ seq)


def test_crc8_vs_reference():
    R = _ref_ext()
    rng = np.random.default_rng(5)
    for n in (0, 1, 9, 10, 37, 255):
        d = rng.integers(0, 256, n, dtype=np.uint8)
        assert ca.bb_crc8(d) == R.ref_crc8(d.ctypes.data, n)


def test_bbframe_parser_restatement_vs_reference():
    """The data-field extraction the GPU tests check against
    (tests/bb_ref.extract) equals the reference's bbframe_parser.cc
    (decode_bbheader + extract_gse_datafield) on valid headers, flipped
    header bits, DFLs that are not byte multiples or too large, the other
    TS/GS values and DFL 0."""
    from tests import bb_ref
    R = _ref_ext()
    rng = np.random.default_rng(8)
    for kbch in (2001, 6051, 7274, 1332):
        for trial in range(120):
            kind = trial % 6
            n_pay = int(rng.integers(0, kbch - 9))
            f = bb_ref.make_bbframe(kbch, rng.integers(0, 256, n_pay, dtype=np.uint8), rng=rng)
            if kind == 1:
                f[rng.integers(0, 10)] ^= 1 << rng.integers(0, 8)  # CRC now fails
            elif kind in (2, 3, 4, 5):
                dfl = {2: 8 * n_pay + int(rng.integers(1, 8)), 3: 8 * kbch - 72, 4: 8 * n_pay, 5: 0}[kind]
                f[4], f[5] = (dfl >> 8) & 0xFF, dfl & 0xFF
                if kind == 4:
                    f[0] = (f[0] & 0x3F) | (int(rng.integers(1, 4)) << 6)
                f[9] = bb_ref.crc8(f[:9])
            ext = bb_ref.extract(f, kbch)
            out = np.zeros(kbch, np.uint8)
            n = R.ref_extract_gse_datafield(f.ctypes.data, kbch, out.ctypes.data, kbch - 10)
            assert n == (len(ext) if ext else 0), (kbch, trial, kind)
            if n:
                assert out[:n].tobytes() == ext


@pytest.mark.parametrize("fs,rate", [(0, "1/4"), (0, "1/2"), (0, "3/4"), (0, "9/10"), (0, "135/180"),
                                     (1, "1/2"), (1, "8/9")])
def test_bench_bbframe_builder_vs_reference(fs, rate):
    """The BBFRAMEs the benchmark sends (stream_bench.bbheader + a full data
    field) equal what the reference transmitter builds (bb_header.cc
    build_gse_bbframe: CCM, roll-off 0.20)."""
    from cusat_amd import stream_bench
    R = _ref_ext()
    kbch = ca.geometry(fs, rate, "QPSK")["k_bch"] // 8
    data = np.random.default_rng(2).integers(0, 256, kbch - 10, dtype=np.uint8)
    out = np.zeros(kbch, np.uint8)
    assert R.ref_build_gse_bbframe(fs, ca.rate_id(rate), data.ctypes.data, kbch - 10, out.ctypes.data) == kbch - 10
    np.testing.assert_array_equal(out[:10], stream_bench.bbheader(kbch))
    np.testing.assert_array_equal(out[10:], data)
