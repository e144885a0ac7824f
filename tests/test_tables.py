"""Code data pins: every LDPC table's canonical text hashes to the
reference's sha256 manifest, the structural invariants of
scripts/check_ldpc_tables.py hold, the schedules match the reference
headers and never pair parity-adjacent layers, and the C++ geometry's
collision plan matches an independent census."""
import hashlib
import json
import os
import re

import pytest

import cusat_amd as ca

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "cusat_amd", "data")


def load_tables():
    tables, cur, lines = {}, None, []
    for line in open(os.path.join(DATA, "ldpc_tables.txt")):
        if line.startswith("#"):
            continue
        if line.strip() == "@table":
            cur, lines = None, []
            continue
        if line.startswith("@links"):
            lt, lmc = map(int, line.split()[1:])
            tables[cur]["canon"] = "".join(lines)
            tables[cur]["links"] = (lt, lmc)
            continue
        lines.append(line)
        if cur is None:
            m = re.match(r"(\w+) N=(\d+) K=(\d+) M=(\d+)", line)
            cur = m.group(1)
            tables[cur] = dict(N=int(m.group(2)), K=int(m.group(3)), M=int(m.group(4)), rows=[])
        else:
            tables[cur]["rows"].append([int(x) for x in line.split()])
    return tables


def load_scheds():
    out, cur = {}, None
    for line in open(os.path.join(DATA, "ldpc_schedules.txt")):
        if line.startswith("#"):
            continue
        if line.startswith("@sched"):
            p = line.split()
            cur = p[1]
            out[cur] = dict(rate=p[2], q=int(p[3][2:]), steps=int(p[4][6:]), pairs=[])
            continue
        a, b = map(int, line.split())
        out[cur]["pairs"].append((a, b))
    return out


TABLES = load_tables()
SCHEDS = load_scheds()


def test_table_checksums_match_reference_manifest():
    manifest = {}
    for line in open(os.path.join(HERE, "golden", "ldpc_table_checksums.txt")):
        if line.strip() and not line.startswith("#"):
            h, name = line.split()
            manifest[name] = h
    assert set(manifest) == set(TABLES)
    for name, t in TABLES.items():
        assert hashlib.sha256(t["canon"].encode()).hexdigest() == manifest[name], name


@pytest.mark.parametrize("name", sorted(TABLES))
def test_table_structure(name):
    """check_ldpc_tables.py validate / links_total / links_max_cn."""
    t = TABLES[name]
    N, K, M = t["N"], t["K"], t["M"]
    R = N - K
    assert M == 360 and R % M == 0 and K == len(t["rows"]) * M
    q = R // M
    cnc = [0] * R
    for row in t["rows"]:
        assert 0 < len(row) <= 255
        for a in row:
            assert 0 <= a < R
        for j in range(M):
            for a in row:
                cnc[(a + j * q) % R] += 1
    lt = M * sum(len(r) for r in t["rows"]) + 2 * R - 1
    assert t["links"] == (lt, max(cnc) + 2)


def test_schedules_match_reference_headers():
    gold = json.load(open(os.path.join(HERE, "golden", "schedule_checksums.json")))
    assert set(gold) == set(SCHEDS)
    for name, s in SCHEDS.items():
        canon = "".join("%d %d\n" % p for p in s["pairs"])
        assert gold[name]["steps"] == s["steps"] == len(s["pairs"])
        assert hashlib.sha256(canon.encode()).hexdigest() == gold[name]["sha256"], name


@pytest.mark.parametrize("name", sorted(SCHEDS))
def test_schedule_invariants(name):
    """Covers every layer once; never pairs layers sharing a parity VN
    (gen_blossom_schedules.py:95-120)."""
    s = SCHEDS[name]
    q = s["q"]
    t = TABLES[name]
    assert q == (t["N"] - t["K"]) // 360
    seen = [l for p in s["pairs"] for l in p if l >= 0]
    assert sorted(seen) == list(range(q))
    for a, b in s["pairs"]:
        if a >= 0 and b >= 0:
            assert (a - b) % q not in (1, q - 1)


def census(name, use_sched=True):
    """Max number of colliding circulants in one step (delta slots)."""
    t = TABLES[name]
    q = (t["N"] - t["K"]) // 360
    layer_rows = {}
    for r, row in enumerate(t["rows"]):
        for a in row:
            layer_rows.setdefault(a % q, []).append(r)
    pairs = SCHEDS[name]["pairs"] if (use_sched and name in SCHEDS) else [(i, -1) for i in range(q)]
    worst = 0
    for a, b in pairs:
        rows = layer_rows.get(a, []) + (layer_rows.get(b, []) if b >= 0 else [])
        cnt = {}
        for r in rows:
            cnt[r] = cnt.get(r, 0) + 1
        worst = max(worst, sum(c for c in cnt.values() if c > 1))
    return worst


def census_plan(name, use_sched=True):
    """Delta slots of the deferred-group plan: a step whose collided rows the
    next step (same iteration) does not touch is applied during that next
    step, so when any step is deferred, even and odd steps get separate slot
    areas (max over even steps + max over odd steps)."""
    t = TABLES[name]
    q = (t["N"] - t["K"]) // 360
    layer_rows = {}
    for r, row in enumerate(t["rows"]):
        for a in row:
            layer_rows.setdefault(a % q, []).append(r)
    pairs = SCHEDS[name]["pairs"] if (use_sched and name in SCHEDS) else [(i, -1) for i in range(q)]
    counts = []
    for a, b in pairs:
        cnt = {}
        for r in layer_rows.get(a, []) + (layer_rows.get(b, []) if b >= 0 else []):
            cnt[r] = cnt.get(r, 0) + 1
        counts.append(cnt)
    slots = [sum(c for c in cnt.values() if c > 1) for cnt in counts]
    defer = [s + 1 < len(counts) and slots[s] > 0 and
             all(r not in counts[s + 1] for r, c in counts[s].items() if c > 1) for s in range(len(counts))]
    if not any(defer):
        return max(slots)
    return max(slots[0::2]) + max(slots[1::2] or [0])


@pytest.mark.parametrize("fs,rate,name", [(0, "1/2", "s2_b4"), (0, "3/4", "s2_b7"),
                                          (0, "2/3", "s2_b6"), (0, "5/6", "s2_b9"),
                                          (0, "9/10", "s2_b11"), (0, "135/180", "s2x_b19"),
                                          (1, "1/2", "s2_c4"), (1, "5/6", "s2_c9")])
def test_geometry_collision_plan(fs, rate, name):
    g = ca.geometry(fs, rate, "QPSK")
    # every listed code's deferred plan fits the LDS (no fallback)
    assert g["n_collision_slots"] == census_plan(name)
    assert census(name) <= g["n_collision_slots"] <= 2 * census(name)
    t = TABLES[name]
    assert g["q"] == (t["N"] - t["K"]) // 360
    # message entries = links_total + 1 (ldpc_decode_kernel.cu:1061-1065)
    assert g["lt_alloc"] == t["links"][0] + 1
    assert g["lanes"] == (2 if name in SCHEDS else 1)
    g_seq = ca.geometry(fs, rate, "QPSK", flags=ca.CTX_SEQUENTIAL_SCHEDULE)
    assert g_seq["lanes"] == 1 and g_seq["n_steps"] == g["q"]
    assert g_seq["n_collision_slots"] == census_plan(name, use_sched=False)


def test_every_ldpc_code_enumerates():
    """The GPU tests that sweep every code see all 51 LDPC codes of the
    tables (34 normal, 17 short), each with its oracle table of equal size."""
    from oracle import oracle_ctypes as oc
    from tests.fec_util import TABLE
    from tests.test_gpu_stream_all_codes import all_codes
    codes = all_codes()
    assert len(codes) == 51
    assert sum(1 for c in codes if c[0] == ca.FECFRAME_SHORT) == 17
    for fs, r, _, _ in codes:
        g = ca.geometry(fs, r, "QPSK")
        c = oc.Code(TABLE[(fs, r)])
        assert (c.N, c.K) == (g["n_ldpc"], g["k_ldpc"]), (fs, r)
