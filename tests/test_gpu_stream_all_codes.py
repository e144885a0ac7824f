"""Every LDPC code of the tables through the persistent stream decoder.

The stream kernel holds one decoder variant per (check degree, exact or
guarded) layout, each compiled as a function of its own, plus the fused BCH
check over every BCH degree (128, 160, 192 normal; 168 short).  One stream
carries frames of all 51 codes (34 normal, 17 short; QPSK, AWGN 2.5 dB above
the Shannon limit of the code rate, so that some frames converge and some
do not), shuffled; every frame's hard bits, trials and BCH flag must equal
the batched decoder of its own code (cusat_ldpc_decode + cusat_bch_check,
which the oracle tests pin), for the 12-wave and the 6-wave workgroups.
"""
import math

import numpy as np
import pytest

import cusat_amd as ca
from tests.test_gpu_stream import build_stream, check, per_code_reference, run_stream

pytestmark = pytest.mark.gpu


def _rate_value(r):
    a, b = r.split("/")
    return int(a) / int(b)


def all_codes():
    """(framesize, rate, "QPSK", Es/N0) for every (framesize, rate) that has
    an LDPC code."""
    out = []
    for fs in (ca.FECFRAME_NORMAL, ca.FECFRAME_SHORT):
        for r in ca.RATES:
            try:
                ca.geometry(fs, r, "QPSK")
            except Exception:
                continue
            eta = 2.0 * _rate_value(r)  # information bits per QPSK symbol
            out.append((fs, r, "QPSK", round(10.0 * math.log10(2.0 ** eta - 1.0) + 2.5, 2)))
    return out


@pytest.mark.parametrize("tuning", [None, {"wg_per_cu": 1, "frames_per_job": 1, "order": 1, "split_edges": 0}])
def test_every_code_in_one_stream(tuning):
    mix = all_codes()
    ctxs, frames, llr, by_code, who, hb = build_stream(mix, 4, 11)
    ref = per_code_reference(ctxs, by_code)
    t = np.concatenate([r[1] for r in ref])
    assert (t < 0).any() and (t >= 0).any(), "the codes need failing and converging frames"
    sc = ca.StreamContext(ctxs, len(frames))
    sc.plan(frames, tuning)
    check(ctxs, frames, who, ref, run_stream(sc, frames, llr, hb))
