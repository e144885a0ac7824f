#!/usr/bin/env python3
"""Regenerate the code data files of this directory from a cuSat source tree.

Two data sets have to be identical to the reference, number for number, and
are therefore extracted rather than retyped:

* ``ldpc_tables.txt`` -- the ETSI EN 302 307-1/-2 accumulator-address tables
  (Annex B/C), written in the canonical text form whose sha256 the reference
  pins in ``lib/cusat_dsp/ldpc_tables/checksums.txt`` (canonical form defined
  at ``scripts/check_ldpc_tables.py:87-95``).  Every table is re-hashed here
  and the run fails on any mismatch.
* ``ldpc_schedules.txt`` -- the 2-lane parallel-layer schedules the reference
  decoder runs for normal frames (``lib/cusat_dsp/cuda/ldpc_schedules/*.h``,
  registry ``ldpc_schedule_registry.h:62-99``).  Layered min-sum results depend
  on the layer order, so bit-exact parity needs the same schedules.
* ``apsk256_points.txt`` -- the explicit-coordinate DVB-S2X 256APSK
  constellations of EN 302 307-2 Table 15d (code ids 20/30 and 22/30), as
  printed in the standard (``lib/cusat_dsp/modulation/apsk256_points.h``);
  the demapper's LLRs depend on every coordinate.

Usage: python extract_reference_data.py /path/to/cusat
"""
import hashlib
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_tables(path):
    text = open(path).read()
    arrays = {}
    for m in re.finditer(
            r"inline constexpr (?:uint8_t|uint16_t|uint32_t) (\w+)\[\] = \{(.*?)\};",
            text, re.S):
        arrays[m.group(1)] = [int(x) for x in m.group(2).replace(",", " ").split()]
    out = []
    for m in re.finditer(
            r"inline constexpr ldpc_code (\w+) = \{\s*"
            r"(\d+), (\d+), (\d+), (\d+), (\d+), (\d+),", text):
        name = m.group(1)
        n, k, mm, lt, lmc, n_rows = (int(m.group(i)) for i in range(2, 8))
        deg = arrays[name + "_deg"]
        off = arrays[name + "_off"]
        addr = arrays[name + "_addr"]
        rows = [addr[off[i]:off[i + 1]] for i in range(n_rows)]
        assert [len(r) for r in rows] == deg
        out.append((name, n, k, mm, lt, lmc, rows))
    return out


def canonical(name, n, k, m, rows):
    parts = ["%s N=%d K=%d M=%d" % (name, n, k, m)]
    parts += [" ".join(str(a) for a in row) for row in rows]
    return "\n".join(parts) + "\n"


def main(ref):
    tdir = os.path.join(ref, "lib", "cusat_dsp", "ldpc_tables")
    manifest = {}
    for line in open(os.path.join(tdir, "checksums.txt")):
        line = line.strip()
        if line and not line.startswith("#"):
            h, name = line.split()
            manifest[name] = h
    tables = parse_tables(os.path.join(tdir, "dvb_s2_tables.h"))
    tables += parse_tables(os.path.join(tdir, "dvb_s2x_tables.h"))
    with open(os.path.join(HERE, "ldpc_tables.txt"), "w") as f:
        f.write("# ETSI EN 302 307-1/-2 LDPC accumulator addresses, one table per block.\n"
                "# Block = canonical text (header line + one address row per line),\n"
                "# sha256(block) equals the reference manifest entry for that table.\n"
                "# The line after each block gives links_total and links_max_cn.\n")
        for name, n, k, m, lt, lmc, rows in tables:
            c = canonical(name, n, k, m, rows)
            h = hashlib.sha256(c.encode()).hexdigest()
            if manifest.get(name) != h:
                raise SystemExit("checksum mismatch for " + name)
            f.write("@table\n" + c + "@links %d %d\n" % (lt, lmc))
    print("tables:", len(tables))

    sdir = os.path.join(ref, "lib", "cusat_dsp", "cuda", "ldpc_schedules")
    reg = open(os.path.join(sdir, "ldpc_schedule_registry.h")).read()
    rates = re.findall(r"case (RATE_\w+): return \{ (\w+), \w+_STEPS, \w+_Q \};", reg)
    with open(os.path.join(HERE, "ldpc_schedules.txt"), "w") as f:
        f.write("# 2-lane parallel-layer schedules for normal FECFRAMEs.\n"
                "# '@sched <table> <rate> q=<q> steps=<n>' then one 'a b' pair per step\n"
                "# (-1 = idle lane).\n")
        for rate, sym in rates:
            text = open(os.path.join(sdir, "ldpc_schedule_%s.h" % sym.lower())).read()
            steps = int(re.search(r"%s_STEPS = (\d+);" % sym, text).group(1))
            q = int(re.search(r"%s_Q\s+= (\d+);" % sym, text).group(1))
            body = re.search(r"%s\[\d+\]\[2\] = \{(.*?)\};" % sym, text, re.S).group(1)
            pairs = re.findall(r"\{\s*(-?\d+),\s*(-?\d+)\}", body)
            assert len(pairs) == steps, sym
            f.write("@sched %s %s q=%d steps=%d\n" % (sym.lower(), rate, q, steps))
            for a, b in pairs:
                f.write("%s %s\n" % (a, b))
    print("schedules:", len(rates))

    text = open(os.path.join(ref, "lib", "cusat_dsp", "modulation", "apsk256_points.h")).read()
    with open(os.path.join(HERE, "apsk256_points.txt"), "w") as f:
        f.write("# EN 302 307-2 Table 15d 256APSK coordinates, unnormalised.\n"
                "# '@points <code id>' then 256 'I Q' lines, line n = label n (MSB first).\n")
        for cid in ("20_30", "22_30"):
            body = re.search(r"APSK256_POINTS_%s\[256\]\[2\] = \{(.*?)\n\};" % cid, text, re.S).group(1)
            pts = re.findall(r"\{\s*(-?\d+\.\d+)f,\s*(-?\d+\.\d+)f\s*\},\s*//\s*([01]{8})", body)
            assert len(pts) == 256 and [int(lab, 2) for _, _, lab in pts] == list(range(256)), cid
            f.write("@points %s\n" % cid)
            for i, q, _ in pts:
                f.write("%s %s\n" % (i, q))
    print("apsk256 tables: 2")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
