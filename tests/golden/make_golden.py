#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from a cuSat source tree.

Run in a container that has the reference tree; the fixtures are data
(checksums, orders, known answers), committed next to this script:

* ldpc_table_checksums.txt  -- the reference's sha256 manifest of the LDPC
  tables (lib/cusat_dsp/ldpc_tables/checksums.txt), unchanged.
* schedule_checksums.json   -- sha256 of each 2-lane schedule's pair list as
  the reference headers hold it (lib/cusat_dsp/cuda/ldpc_schedules/*.h),
  canonical text "a b\\n" per step.
* kat.json                  -- known answers of the reference's unit tests:
  message update (qa_ldpc_message_update.cc:10-27), acceptance truth table
  (qa_rx_fec_chain.cc:9-17), BCH degrees / g1 (qa_bch_code.cc:36-55),
  interleaver orders (qa_bit_interleaver.cc), 32APSK ring ratios
  (qa_constellation.cc:102-128).

Usage: python tests/golden/make_golden.py [/path/to/cusat]
"""
import hashlib
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    shutil.copyfile(os.path.join(ref, "lib/cusat_dsp/ldpc_tables/checksums.txt"),
                    os.path.join(HERE, "ldpc_table_checksums.txt"))
    sdir = os.path.join(ref, "lib/cusat_dsp/cuda/ldpc_schedules")
    out = {}
    for fn in sorted(os.listdir(sdir)):
        m = re.match(r"ldpc_schedule_(s2x?_b\d+)\.h$", fn)
        if not m:
            continue
        sym = m.group(1).upper()
        text = open(os.path.join(sdir, fn)).read()
        body = re.search(r"%s\[\d+\]\[2\] = \{(.*?)\};" % sym, text, re.S).group(1)
        pairs = re.findall(r"\{\s*(-?\d+),\s*(-?\d+)\}", body)
        canon = "".join("%s %s\n" % p for p in pairs)
        out[m.group(1)] = {"steps": len(pairs),
                           "sha256": hashlib.sha256(canon.encode()).hexdigest()}
    json.dump(out, open(os.path.join(HERE, "schedule_checksums.json"), "w"), indent=1,
              sort_keys=True)

    kat = {
        # (raw new message, old message) -> (stored message, delta)
        "message_update": [[300, 127, 127, 0], [300, -127, 127, 254],
                           [-300, 127, -127, -254], [42, -5, 42, 47]],
        # (trials remaining, bch valid) -> accepted
        "acceptance": [[25, 1, 1], [0, 1, 1], [-1, 1, 0], [3, 0, 0], [-1, 0, 0]],
        # (framesize, n_bits, t) -> (parity bits, k bits or null)
        "bch_degree": [[0, 16200, 12, 192, 16008], [0, 48600, 10, 160, None],
                       [0, 57600, 8, 128, None], [1, 3240, 12, 168, None]],
        "bch_g1_terms": {"normal": [0, 2, 3, 5, 16], "short": [0, 1, 3, 5, 14]},
        # (modulation, rate, framesize) -> column order, rows
        "interleaver": [["8PSK", "2/3", 0, [0, 1, 2], 21600], ["8PSK", "3/5", 0, [2, 1, 0], None],
                        ["32APSK", "3/4", 1, None, 3240],
                        ["32APSK", "2/3", 0, [2, 1, 4, 3, 0], None],
                        ["64APSK", "4/5", 0, [1, 2, 4, 0, 5, 3], None],
                        ["256APSK", "135/180", 0, [5, 0, 7, 4, 3, 6, 1, 2], None],
                        ["128APSK", "135/180", 0, [4, 2, 5, 0, 3, 1, 6], 9258]],
        "apsk32_ratios": [["3/4", 2.84, 5.27], ["4/5", 2.72, 4.87], ["5/6", 2.64, 4.64],
                          ["8/9", 2.54, 4.33], ["9/10", 2.53, 4.30]],
        "apsk16_ratio_3_4": 2.85,
    }
    json.dump(kat, open(os.path.join(HERE, "kat.json"), "w"), indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
