"""Baseband reference for the tests, written from EN 302 307-1: the BB
scrambler register of clause 5.2.2 (figure 5), BBHEADER fields of table 3
and its CRC-8 (clause 5.1.6), restated bit by bit."""
import numpy as np


def prbs_bits(n):
    """Stages 1..15, initial content 100101010000000, output = stage 14 XOR
    stage 15, shifted in at stage 1."""
    st = [0] + [int(c) for c in "100101010000000"]
    out = np.zeros(n, np.uint8)
    for k in range(n):
        b = st[14] ^ st[15]
        st[2:16] = st[1:15]
        st[1] = b
        out[k] = b
    return out


def crc8(data):
    r = 0
    for byte in bytes(data):
        r ^= byte
        for _ in range(8):
            r = ((r << 1) ^ 0xD5) & 0xFF if r & 0x80 else (r << 1) & 0xFF
    return r


def make_bbframe(kbch, payload, ts_gs=0, dfl_bits=None, rng=None, corrupt_bit=None):
    """BBHEADER (MATYPE-1 SIS=1, CCM=1, RO=2 | TS/GS) + payload + random
    padding, NOT scrambled."""
    rng = rng or np.random.default_rng(0)
    f = rng.integers(0, 256, kbch, dtype=np.uint8)
    dfl = 8 * len(payload) if dfl_bits is None else dfl_bits
    f[0] = (ts_gs << 6) | (1 << 5) | (1 << 4) | 2
    f[1] = 0
    f[2] = f[3] = 0
    f[4], f[5] = dfl >> 8, dfl & 0xFF
    f[6] = 0
    f[7] = f[8] = 0xFF
    f[9] = crc8(f[:9])
    f[10:10 + len(payload)] = payload
    if corrupt_bit is not None:
        f[corrupt_bit // 8] ^= 0x80 >> (corrupt_bit % 8)
    return f


def extract(frame_descrambled, kbch):
    """Data-field bytes of a descrambled BBFRAME, or None (the checks of
    table 3: CRC over 10 bytes == 0, TS/GS generic packetized, DFL a byte
    multiple within the frame)."""
    h = frame_descrambled[:10]
    if crc8(h) != 0 or (h[0] >> 6) != 0:
        return None
    dfl = (int(h[4]) << 8) | int(h[5])
    if dfl % 8 or dfl > 8 * kbch - 80:
        return None
    return bytes(frame_descrambled[10:10 + dfl // 8])
