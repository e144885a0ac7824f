"""Pins of the CPU oracle (oracle/) against the reference's known answers
and against what the mathematics fixes, independently of the oracle's own
formulas."""
import json
import math
import os

import numpy as np
import pytest

import cusat_amd as ca
from oracle import oracle_ctypes as oc
from tests.test_tables import SCHEDS, TABLES

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))


# ---------------------------------------------------------------- messages

def test_message_update_kats():
    """qa_ldpc_message_update.cc:10-27: stored message = clamp(raw), delta =
    stored - old."""
    for raw, old, msg, delta in KAT["message_update"]:
        m = oc.clamp_msg(raw)
        assert m == msg and m - old == delta


def _check_node_brute(x, beta, alpha):
    """Extrinsic min-sum by definition: for edge e, the minimum over the
    OTHER edges of max(|x| - beta, 0) (|-32768| = 32767), with the sign
    parity of the other edges (x < 0 counts as 1), scaled (alpha/64,
    round half up on the magnitude) and clamped to +-127."""
    x = [int(v) for v in x]
    mags = [max(min(abs(v), 32767) - beta, 0) for v in x]
    out = []
    for e in range(len(x)):
        others = [mags[k] for k in range(len(x)) if k != e]
        neg = sum(1 for k in range(len(x)) if k != e and x[k] < 0) & 1
        m = min(others)
        if alpha:
            m = (m * alpha + 32) >> 6
        m = min(m, 127)
        out.append(-m if neg else m)
    return out


@pytest.mark.parametrize("seed", range(12))
def test_check_node_vs_definition(seed):
    rng = np.random.default_rng(seed)
    for _ in range(60):
        deg = int(rng.integers(3, 33))
        kind = rng.integers(0, 4)
        if kind == 0:
            x = rng.integers(-40, 41, deg)
        elif kind == 1:
            x = rng.integers(-32768, 32768, deg)
        elif kind == 2:
            x = rng.choice([-32768, 32767, 0, 1, -1, 2], deg)
        else:
            x = rng.integers(-3, 4, deg)  # many ties
        beta = int(rng.choice([0, 1, 2, 5]))
        alpha = int(rng.choice([0, 0, 48, 64, 80]))
        got = oc.check_node(x.astype(np.int16), beta, alpha)
        np.testing.assert_array_equal(got, _check_node_brute(x, beta, alpha))


def test_acceptance_truth_table():
    """qa_rx_fec_chain.cc:9-17 -- accepted iff trials >= 0 and BCH valid."""
    for trials, valid, acc in KAT["acceptance"]:
        assert int(trials >= 0 and bool(valid)) == acc


# --------------------------------------------------------------------- BCH

def _gf_mulmod(a, b, mod, deg):
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if (a >> deg) & 1:
            a ^= mod
    return r


def _poly_eval_gf(coeffs_bits, x, mod, deg):
    """Evaluate the GF(2)[z] polynomial with bit-coefficients at x in
    GF(2^deg) = GF(2)[a]/mod (Horner)."""
    acc = 0
    for i in range(coeffs_bits.bit_length() - 1, -1, -1):
        acc = _gf_mulmod(acc, x, mod, deg)
        if (coeffs_bits >> i) & 1:
            acc ^= 1
    return acc


def _g_int(g):
    return sum(int(w) << (64 * k) for k, w in enumerate(g))


@pytest.mark.parametrize("fs,n,t,p,k", [tuple(v) for v in KAT["bch_degree"]])
def test_bch_degrees(fs, n, t, p, k):
    d, _ = oc.bch_generator(fs, t)
    assert d == p
    if k is not None:
        assert n - d == k


@pytest.mark.parametrize("fs,m,t", [(0, 16, 12), (0, 16, 10), (0, 16, 8), (1, 14, 12)])
def test_bch_generator_has_the_2t_consecutive_roots(fs, m, t):
    """g1 is the field polynomial (KAT: its terms); in GF(2^m) = GF(2)[a]/g1
    the generator must vanish at a, a^2, ..., a^2t -- the defining property
    of a t-error-correcting BCH code -- and have degree m t."""
    terms = KAT["bch_g1_terms"]["normal" if fs == 0 else "short"]
    g1 = sum(1 << e for e in terms)
    _, g1_or = oc.bch_generator(fs, 1)
    assert _g_int(g1_or) == g1
    d, g = oc.bch_generator(fs, t)
    gi = _g_int(g)
    assert d == m * t and gi & 1
    alpha = 2  # the class of z
    x = 1
    for j in range(1, 2 * t + 1):
        x = _gf_mulmod(x, alpha, g1, m)
        assert _poly_eval_gf(gi, x, g1, m) == 0, j


def test_bch_codewords_linearity_and_detection():
    rng = np.random.default_rng(3)
    for fs, rate, t in [(0, "3/4", 12), (0, "2/3", 10), (0, "9/10", 8), (1, "1/2", 12)]:
        g = ca.geometry(fs, rate, "QPSK")
        msg = rng.integers(0, 256, (3, g["k_bch"] // 8), dtype=np.uint8)
        cw = ca.bch_encode(fs, rate, msg)
        n = g["n_bch"]
        assert np.array_equal(cw[:, : g["k_bch"] // 8], msg)  # systematic
        assert list(oc.bch_check(cw, n, fs, t)) == [1, 1, 1]
        assert oc.bch_check(cw[0] ^ cw[1], n, fs, t)[0] == 1  # linear
        bad = cw.copy()
        bad[0, 7] ^= 0x10
        bad[1, -1] ^= 0x01
        assert list(oc.bch_check(bad, n, fs, t)) == [0, 0, 1]


# ------------------------------------------------------------- exp / log

def test_exp_log_error_bounds():
    xs = np.linspace(-15.999, 0.0, 4001)
    for x in xs:
        e = oc.exp_neg(float(np.float32(x)))
        ref = math.exp(float(np.float32(x)))
        assert abs(e - ref) <= 4e-7 * ref
    for v in np.concatenate([np.linspace(1.0, 256.0, 3001), np.float32([1e-30, 1e-3, 0.5])]):
        v = float(np.float32(v))
        assert abs(oc.log_pos(v) - math.log(v)) <= 6e-7 * max(1.0, abs(math.log(v)))


# ---------------------------------------------------------------- demapper

def _maxlog_psk8(re, im, scale):
    pts = [(math.cos(a), math.sin(a)) for a in
           [math.pi / 4, 0, math.pi, 5 * math.pi / 4, math.pi / 2, -math.pi / 4, 3 * math.pi / 4,
            -math.pi / 2]]
    out = []
    for b in range(3):
        d0 = min((re - p[0]) ** 2 + (im - p[1]) ** 2 for i, p in enumerate(pts) if not (i >> (2 - b)) & 1)
        d1 = min((re - p[0]) ** 2 + (im - p[1]) ** 2 for i, p in enumerate(pts) if (i >> (2 - b)) & 1)
        out.append((d1 - d0) * scale)
    return out


def test_8psk_closed_form():
    """8PSK (demapper_kernel.cu:76-107): derotate by e^{-j pi/8}, then
    llr0 = (|cr| - |ci|)/sqrt(2) * s, llr1 = cr s, llr2 = ci s with
    s = 2 sin(pi/8) gain.  Pins: (a) float64 re-evaluation of that formula
    within one count; (b) every sign agrees with brute-force max-log over the
    Gray label map of figure 10 (positive = bit 0); (c) the first bit is the
    exact max-log LLR at the FACTOR=2 scale (gain / 2)."""
    rng = np.random.default_rng(1)
    sym = (rng.normal(0, 0.7, (1, 3000, 2))).astype(np.float32)
    gain = 23.5
    ps = float(np.float32(2.0 * np.float32(0.38268343236508977)) * np.float32(gain))
    col = np.array([0, 3000, 6000, 0, 0, 0, 0, 0], np.uint32)
    llr = oc.demap(sym, 3, None, col, 9000, 0.0, ps, gain)[0]
    c8, s8 = math.cos(math.pi / 8), math.sin(math.pi / 8)
    for j in range(3000):
        re, im = float(sym[0, j, 0]), float(sym[0, j, 1])
        cr, ci = re * c8 + im * s8, -re * s8 + im * c8
        f = [(abs(cr) - abs(ci)) / math.sqrt(2) * ps, cr * ps, ci * ps]
        ml = _maxlog_psk8(re, im, gain / 2.0)
        for b in range(3):
            got = int(llr[col[b] + j])
            assert abs(got - f[b]) <= 1.0 + 1e-6
            if abs(ml[b]) > 1.0:
                assert (got > 0) == (ml[b] > 0)
        assert abs(int(llr[j]) - ml[0]) <= 1.0 + 1e-6


def test_qpsk_closed_form():
    rng = np.random.default_rng(2)
    sym = rng.normal(0, 1.5, (1, 500, 2)).astype(np.float32)
    qs = float(np.float32(2 * 1.4142135623730951) / np.float32(0.5))
    llr = oc.demap(sym, 2, None, np.array([0, 1, 0, 0, 0, 0, 0, 0], np.uint32), 1000, qs, 0, 0)[0]
    ref = np.clip(np.rint(sym[0].reshape(-1).astype(np.float64) * qs), -32768, 32767)
    assert np.max(np.abs(llr - ref)) <= 1


@pytest.mark.parametrize("mod,rate", [("16APSK", "2/3"), ("32APSK", "3/4"), ("64APSK", "4/5"),
                                      ("128APSK", "135/180"), ("256APSK", "135/180")])
def test_apsk_logmap_vs_float64(mod, rate):
    """Two-pass log-MAP (demapper_kernel.cu:108-183) against float64
    evaluation of the same expression with libm exp/log: within one count."""
    pts, col, nm, rows = ca.modcod_constellation(0, rate, mod)
    P = pts.reshape(-1, 2).astype(np.float64)
    rng = np.random.default_rng(7)
    L = 200
    lab = rng.integers(0, 1 << nm, L)
    sym = (P[lab] + rng.normal(0, 0.15, (L, 2))).astype(np.float32)[None]
    gain = 30.0
    c = np.arange(8, dtype=np.uint32) * L
    llr = oc.demap(sym, nm, pts, c, L * nm, 0, 0, gain)[0]
    for j in range(L):
        d = ((sym[0, j].astype(np.float64) - P) ** 2).sum(1)
        for b in range(nm):
            m1 = (np.arange(len(P)) >> (nm - 1 - b)) & 1
            d0, d1 = d[m1 == 0], d[m1 == 1]
            s0 = np.sum(np.exp(-(d0 - d0.min()) * gain / 4)[(d0 - d0.min()) * gain / 4 < 16])
            s1 = np.sum(np.exp(-(d1 - d1.min()) * gain / 4)[(d1 - d1.min()) * gain / 4 < 16])
            ref = (d1.min() - d0.min()) * gain / 2 + 2 * (math.log(s0) - math.log(s1))
            assert abs(int(llr[c[b] + j]) - ref) <= 1.0 + 1e-3, (j, b)


# ----------------------------------------------------------------- decoder

def _py_layered(name, llr, T, use_sched=True):
    """Independent restatement of the layered decoder in the reference's own
    data layout: per-check edge lists built in the append order of
    ldpc_decode_kernel.cu:1006-1019 ((a + jj q) mod R for bit 360 r + jj),
    plain Python integers.  Serialisation: all reads of a step first; data
    writes applied lane by lane, circulant by circulant (table order)."""
    t = TABLES[name]
    N, K = t["N"], t["K"]
    R = N - K
    q = R // 360
    edges = [[] for _ in range(R)]  # (vn, rank of the circulant in its layer)
    rank = {}
    per_layer = {}
    for r, row in enumerate(t["rows"]):
        for n, a in enumerate(row):
            per_layer.setdefault(a % q, []).append((r, n))
    for i, lst in per_layer.items():
        for k, rn in enumerate(lst):
            rank[rn] = k
    for r, row in enumerate(t["rows"]):
        for jj in range(360):
            for n, a in enumerate(row):
                edges[(a + jj * q) % R].append((360 * r + jj, rank[(r, n)]))
    if use_sched and name in SCHEDS:
        steps = SCHEDS[name]["pairs"]
    else:
        steps = [(i, -1) for i in range(q)]

    def sat(v):
        return max(-32768, min(32767, v))

    app_d = [max(int(v), -32767) for v in llr[:K]]
    app_p = [max(int(v), -32767) for v in llr[K:]]
    msg = {}
    res = -1
    for it in range(T + 1):
        ok = True
        for c in range(R):
            neg = (app_p[c] < 0) ^ (c > 0 and app_p[c - 1] < 0)
            for v, _ in edges[c]:
                neg ^= app_d[v] < 0
            if neg:
                ok = False
                break
        if ok:
            res = T - it
            break
        if it == T:
            break
        for step in steps:
            writes = []
            pwrites = []
            for lane, i in enumerate(step):
                if i < 0:
                    continue
                for j in range(360):
                    c = q * j + i
                    vals = [app_d[v] for v, _ in edges[c]] + [app_p[c]] + ([app_p[c - 1]] if c > 0 else [])
                    old = msg.get(c, [0] * len(vals))
                    x = [sat(a - o) for a, o in zip(vals, old)]
                    y = _check_node_brute(x, 1, 0)
                    msg[c] = y
                    ne = len(edges[c])
                    for e, (v, k) in enumerate(edges[c]):
                        writes.append((lane, k, j, v, y[e] - old[e]))
                    pwrites.append((c, sat(x[ne] + y[ne])))
                    if c > 0:
                        pwrites.append((c - 1, sat(x[ne + 1] + y[ne + 1])))
            for lane, k, j, v, d in sorted(writes, key=lambda w: (w[0], w[1], w[2])):
                app_d[v] = sat(app_d[v] + d)
            for c, val in pwrites:
                app_p[c] = val
    hard = np.packbits(np.array([1 if v < 0 else 0 for v in app_d], np.uint8))
    return hard, res


@pytest.mark.parametrize("name,fs,rate,es,T,use_sched", [
    ("s2_c4", 1, "1/2", 1.0, 6, True),     # short: sequential, intra-layer collisions
    ("s2_c9", 1, "5/6", 8.5, 5, True),     # short 5/6: weight-2/3 circulants
    ("s2_b9", 0, "5/6", 6.6, 2, True),     # normal 5/6: lane-pair collisions
])
def test_oracle_vs_independent_restatement(name, fs, rate, es, T, use_sched):
    code = oc.Code(name, use_sched)
    sym, _, _ = ca.gen_frames(fs, rate, "QPSK", 17, 1, es)
    snr = np.float32(10.0) ** (np.float32(es) / np.float32(10.0))
    qs = float(np.float32(2 * np.float32(1.4142135623730951)) / (np.float32(1.0) / snr))
    llr = oc.demap(sym, 2, None, np.array([0, 1, 0, 0, 0, 0, 0, 0], np.uint32), code.N, qs, 0, 0)
    h_o, r_o = code.decode(llr, T)
    h_p, r_p = _py_layered(name, llr[0], T, use_sched)
    assert r_o[0] == r_p
    np.testing.assert_array_equal(h_o[0], h_p)


def _np_layered(name, llr, T, use_sched=True):
    """A second independent restatement, vectorised with numpy so that
    normal frames run many iterations in seconds.  The same semantics as
    _py_layered, formulated differently: per layer an edge matrix [360,
    cnt] (checks q j + i, edges in the append order (r, jj, address index)
    of ldpc_decode_kernel.cu:1006-1019); the check node as the extrinsic
    minimum from the two smallest magnitudes of a sort, the sign from the
    count of negative inputs among the other edges; data writes applied per
    (lane, circulant) (the 360 VNs of one circulant are distinct), then
    the parity writes from the read-phase values."""
    t = TABLES[name]
    N, K = t["N"], t["K"]
    R = N - K
    q = R // 360
    rows = t["rows"]
    # every edge: check, vn, (r, jj, n) for the append order, circulant rank
    layer_circ = {}
    for r, row in enumerate(rows):
        for n, a in enumerate(row):
            layer_circ.setdefault(a % q, []).append((r, n, a))
    E, RK = {}, {}
    jj = np.arange(360)
    for i in range(q):
        circs = layer_circ.get(i, [])
        if not circs:
            E[i] = np.zeros((360, 0), np.int64)
            RK[i] = np.zeros((360, 0), np.int64)
            continue
        # check q j + i receives from circulant (r, n, a) the bit 360 r + jj
        # with (a + jj q) mod R = q j + i
        vn = np.zeros((360, len(circs)), np.int64)
        key = np.zeros((360, len(circs)), np.int64)
        for k, (r, n, a) in enumerate(circs):
            j_of_jj = ((a + jj * q) % R - i) // q
            vn[j_of_jj, k] = 360 * r + jj
            key[j_of_jj, k] = (r * 360 + jj) * 64 + n
        order = np.argsort(key, axis=1, kind="stable")
        E[i] = np.take_along_axis(vn, order, 1)
        RK[i] = order  # rank of the circulant (layer order) of each edge
    if use_sched and name in SCHEDS:
        steps = [tuple(p) for p in SCHEDS[name]["pairs"]]
    else:
        steps = [(i, -1) for i in range(q)]
    jq = np.arange(360) * q

    def sat(v):
        return np.clip(v, -32768, 32767)

    app_d = np.maximum(llr[:K].astype(np.int64), -32767)
    app_p = np.maximum(llr[K:].astype(np.int64), -32767)
    msg = {}
    res = -1
    for it in range(T + 1):
        ok = True
        for i in range(q):
            c = jq + i
            neg = (app_d[E[i]] < 0).sum(1) + (app_p[c] < 0)
            neg = neg + np.where(c > 0, app_p[np.maximum(c - 1, 0)] < 0, False)
            if (neg & 1).any():
                ok = False
                break
        if ok:
            res = T - it
            break
        if it == T:
            break
        for step in steps:
            reads = []
            for i in step:
                if i < 0:
                    continue
                c = jq + i
                cols = [app_d[E[i]], app_p[c][:, None]]
                has2 = c > 0
                cols.append(np.where(has2, app_p[np.maximum(c - 1, 0)], 0)[:, None])
                vals = np.concatenate(cols, 1)
                old = msg.get(i, np.zeros_like(vals))
                x = sat(vals - old)
                mag = np.maximum(np.minimum(np.abs(x), 32767) - 1, 0)
                neg = x < 0
                # check (0, 0) has no second parity edge: out of the minimum
                # and the sign
                if i == 0:
                    mag[0, -1] = 1 << 20
                    neg[0, -1] = False
                srt = np.sort(mag, 1)
                m = np.where(mag == srt[:, :1], srt[:, 1:2], srt[:, :1])
                sgn = (neg.sum(1, keepdims=True) - neg) & 1
                y = np.clip(np.where(sgn == 1, -m, m), -127, 127)
                if i == 0:
                    y[0, -1] = 0
                msg[i] = y
                reads.append((i, c, has2, x, y, old))
            for i, c, has2, x, y, old in reads:
                ne = E[i].shape[1]
                d = y[:, :ne] - old[:, :ne]
                for k in range(ne):
                    cols = np.nonzero(RK[i] == k)
                    v = E[i][cols]
                    app_d[v] = sat(app_d[v] + d[cols])
            for i, c, has2, x, y, old in reads:
                ne = E[i].shape[1]
                app_p[c] = sat(x[:, ne] + y[:, ne])
                cc = c[has2] - 1
                app_p[cc] = sat(x[has2, ne + 1] + y[has2, ne + 1])
    hard = np.packbits((app_d < 0).astype(np.uint8))
    return hard, res


def test_vectorised_restatement_equals_python_restatement():
    """The two independent restatements agree (short codes, noisy frames)."""
    for name, fs, rate, es, T in [("s2_c4", 1, "1/2", 1.0, 6), ("s2_c9", 1, "5/6", 8.5, 4),
                                  ("s2_b9", 0, "5/6", 6.6, 2)]:
        code = oc.Code(name)
        sym, _, _ = ca.gen_frames(fs, rate, "QPSK", 17, 1, es)
        snr = np.float32(10.0) ** (np.float32(es) / np.float32(10.0))
        qs = float(np.float32(2 * np.float32(1.4142135623730951)) / (np.float32(1.0) / snr))
        llr = oc.demap(sym, 2, None, np.array([0, 1, 0, 0, 0, 0, 0, 0], np.uint32), code.N, qs, 0, 0)[0]
        h_p, r_p = _py_layered(name, llr, T)
        h_n, r_n = _np_layered(name, llr, T)
        assert r_p == r_n, name
        np.testing.assert_array_equal(h_p, h_n)


def _rail_frame(fs, rate, N, seed, K=None):
    """A codeword whose bits sit either at the int16 rails (+32767, -32768
    and just inside, correct signs) or at small magnitudes with a few of
    those signs wrong (3 % x (1 - K/N) of the small ones): the decoder
    corrects them while the APP sums and the parity writes of the railed
    bits saturate."""
    _, _, cw = ca.gen_frames(fs, rate, "QPSK", seed, 1, 1000.0, want_cw=True)
    bits = np.unpackbits(cw[0])[:N]
    rng = np.random.default_rng(seed)
    big = rng.random(N) < 0.7
    mag = np.where(big, rng.choice([32768, 32767, 32700], N), rng.integers(4, 41, N))
    wrong = ~big & (rng.random(N) < 0.03 * (1.0 - (K or N // 2) / N))
    s = np.where(bits == 0, 1, -1) * np.where(wrong, -1, 1)
    return np.clip(s * mag, -32768, 32767).astype(np.int16)


@pytest.mark.parametrize("name,fs,rate,th,T", [
    ("s2_b4", 0, "1/2", 1.80, 12),   # config 1's code (QPSK 1/2 normal)
    ("s2_b7", 0, "3/4", 4.43, 12),   # config 2's code (3/4 normal)
    ("s2_b6", 0, "2/3", 3.50, 12),   # config 4's code (2/3 normal, t = 10)
    ("s2_b11", 0, "9/10", 6.72, 10),  # 9/10: lane-pair collisions
])
def test_oracle_vs_vectorised_restatement_normal_frames(name, fs, rate, th, T):
    """The oracle's hard bits and remaining-trial count equal the vectorised
    restatement on normal frames of the configs' codes with T >= 10: a
    QPSK frame at the MODCOD's threshold + 0.5 dB (converges after a few
    iterations), one at threshold - 0.4 dB (fails after T), and a frame at
    the int16 rails."""
    code = oc.Code(name)
    frames = []
    for es, seed in ((th + 0.5, 29), (th - 0.4, 30)):
        sym, _, _ = ca.gen_frames(fs, rate, "QPSK", seed, 1, es)
        snr = np.float32(10.0) ** (np.float32(es) / np.float32(10.0))
        qs = float(np.float32(2 * np.float32(1.4142135623730951)) / (np.float32(1.0) / snr))
        frames.append(oc.demap(sym, 2, None, np.array([0, 1, 0, 0, 0, 0, 0, 0], np.uint32), code.N, qs, 0, 0)[0])
    frames.append(_rail_frame(fs, rate, code.N, 31, code.K))
    results = []
    for f in frames:
        h_o, r_o = code.decode(f[None], T)
        h_n, r_n = _np_layered(name, f, T)
        assert r_o[0] == r_n
        np.testing.assert_array_equal(h_o[0], h_n)
        results.append(r_n)
    assert 0 <= results[0] < T and results[1] == -1 and 0 <= results[2] < T, results


@pytest.mark.parametrize("fs,rate,name", [(0, r, n) for r, n in
                                          [("1/4", "s2_b1"), ("1/2", "s2_b4"), ("3/4", "s2_b7"),
                                           ("9/10", "s2_b11")]] +
                         [(1, "2/3", "s2_c6"), (1, "8/9", "s2_c10")])
def test_noiseless_codewords_converge_at_trial_zero(fs, rate, name):
    """qa_ldpc_encode.cc:68-86 / bench_ldpc_decode.cc:171-193: a noiseless
    codeword (any magnitude) passes the syndrome before iteration 0 and
    positive LLR means bit 0."""
    code = oc.Code(name)
    _, info, cw = ca.gen_frames(fs, rate, "QPSK", 42, 2, 1000.0, want_cw=True)
    bits = np.unpackbits(cw, axis=1)[:, : code.N]
    for mag in (127, 3000, 32767):
        llr = np.where(bits == 0, mag, -mag).astype(np.int16)
        hard, res = code.decode(llr, 50)
        assert list(res) == [50, 50]
        np.testing.assert_array_equal(hard, cw[:, : code.K // 8])


def test_result_range_and_determinism():
    code = oc.Code("s2_b7")
    rng = np.random.default_rng(0)
    llr = rng.integers(-20, 21, (3, code.N)).astype(np.int16)
    h1, r1 = code.decode(llr, 7, nthreads=1)
    h2, r2 = code.decode(llr, 7, nthreads=3)
    assert np.array_equal(h1, h2) and np.array_equal(r1, r2)
    assert all(r == -1 or 0 <= r <= 7 for r in r1)
