"""ctypes access to the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by bench.py's CPU-baseline leg and by
__graft_entry__.smoke(), always as the checker.  The product (cusat_amd/)
never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liboracle.so")
TABLES = os.path.join(ROOT, "cusat_amd", "data", "ldpc_tables.txt")
SCHEDS = os.path.join(ROOT, "cusat_amd", "data", "ldpc_schedules.txt")

_lib = None

TABLE = {  # (framesize, rate) -> oracle table name
    (0, "1/4"): "s2_b1", (0, "1/3"): "s2_b2", (0, "2/5"): "s2_b3", (0, "1/2"): "s2_b4",
    (0, "3/5"): "s2_b5", (0, "2/3"): "s2_b6", (0, "3/4"): "s2_b7", (0, "4/5"): "s2_b8",
    (0, "5/6"): "s2_b9", (0, "8/9"): "s2_b10", (0, "9/10"): "s2_b11",
    (0, "135/180"): "s2x_b19", (0, "128/180"): "s2x_b17", (0, "154/180"): "s2x_b21",
    (0, "96/180"): "s2x_b12", (0, "22/30"): "s2x_b24", (0, "7/9"): "s2x_b10", (1, "7/15"): "s2x_c4",
    (1, "1/4"): "s2_c1", (1, "1/3"): "s2_c2", (1, "2/5"): "s2_c3", (1, "1/2"): "s2_c4",
    (1, "3/5"): "s2_c5", (1, "2/3"): "s2_c6", (1, "3/4"): "s2_c7", (1, "4/5"): "s2_c8",
    (1, "5/6"): "s2_c9", (1, "8/9"): "s2_c10",
    # the other DVB-S2X codes, by their table in EN 302 307-2 annexes B and C
    (0, "13/45"): "s2x_b2", (0, "9/20"): "s2x_b3", (0, "11/20"): "s2x_b4", (0, "26/45"): "s2x_b5",
    (0, "28/45"): "s2x_b6", (0, "23/36"): "s2x_b7", (0, "25/36"): "s2x_b8", (0, "13/18"): "s2x_b9",
    (0, "90/180"): "s2x_b11", (0, "100/180"): "s2x_b13", (0, "104/180"): "s2x_b14",
    (0, "116/180"): "s2x_b15", (0, "124/180"): "s2x_b16", (0, "132/180"): "s2x_b18",
    (0, "140/180"): "s2x_b20", (0, "18/30"): "s2x_b22", (0, "20/30"): "s2x_b23",
    (1, "11/45"): "s2x_c1", (1, "4/15"): "s2x_c2", (1, "14/45"): "s2x_c3", (1, "8/15"): "s2x_c5",
    (1, "26/45"): "s2x_c6", (1, "32/45"): "s2x_c7",
}


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i = ctypes.c_void_p, ctypes.c_int
        L.oc_code_open.restype = vp
        L.oc_code_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i]
        L.oc_code_close.argtypes = [vp]
        L.oc_code_info.argtypes = [vp] + [ctypes.POINTER(i)] * 6
        L.oc_ldpc_decode.argtypes = [vp, vp, i, vp, vp, i, vp, vp]
        L.oc_ldpc_decode_batch.argtypes = [vp, vp, i, i, vp, vp, i, vp, vp, i]
        L.oc_ldpc_syndrome_ok.argtypes = [vp, vp]
        L.oc_check_node.argtypes = [vp, i, i, i, vp]
        L.oc_clamp_msg.argtypes = [i]
        L.oc_bch_generator.argtypes = [i, i, vp]
        L.oc_bch_check.argtypes = [vp, i, i, i]
        L.oc_demap.argtypes = [vp, i, i, vp, i, vp, i, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, vp]
        L.oc_exp_neg.restype = ctypes.c_float
        L.oc_exp_neg.argtypes = [ctypes.c_float]
        L.oc_log_pos.restype = ctypes.c_float
        L.oc_log_pos.argtypes = [ctypes.c_float]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Code:
    """One LDPC code of the oracle (table name, e.g. 's2_b7')."""

    def __init__(self, name, use_sched=True):
        self.name = name
        self._h = lib().oc_code_open(TABLES.encode(), SCHEDS.encode(), name.encode(),
                                     1 if use_sched else 0)
        if not self._h:
            raise ValueError("oracle: unknown table " + name)
        v = [ctypes.c_int() for _ in range(6)]
        lib().oc_code_info(self._h, *[ctypes.byref(x) for x in v])
        self.N, self.K, self.q, self.n_steps, self.cnl, self.lt_alloc = [x.value for x in v]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oc_code_close(self._h)
            self._h = None

    def decode(self, llr, max_trials, beta=None, alpha=None, n_param_iters=0, nthreads=8):
        llr = np.ascontiguousarray(llr, np.int16).reshape(-1, self.N)
        B = llr.shape[0]
        hard = np.zeros((B, self.K // 8), np.uint8)
        res = np.zeros(B, np.int32)
        if beta is not None:
            beta = np.ascontiguousarray(beta, np.int16)
        if alpha is not None:
            alpha = np.ascontiguousarray(alpha, np.int16)
        rc = lib().oc_ldpc_decode_batch(self._h, _p(llr), B, max_trials, _p(beta), _p(alpha),
                                        n_param_iters, _p(hard), _p(res), nthreads)
        assert rc == 0
        return hard, res

    def syndrome_ok(self, llr):
        llr = np.ascontiguousarray(llr, np.int16)
        return bool(lib().oc_ldpc_syndrome_ok(self._h, _p(llr)))


def check_node(x, beta=1, alpha64=0):
    x = np.ascontiguousarray(x, np.int16)
    out = np.zeros_like(x)
    lib().oc_check_node(_p(x), len(x), beta, alpha64, _p(out))
    return out


def clamp_msg(v):
    return lib().oc_clamp_msg(int(v))


def bch_generator(framesize, t):
    g = np.zeros(4, np.uint64)
    d = lib().oc_bch_generator(framesize, t, _p(g))
    return d, g


def bch_check(frames, nbits, framesize, t):
    frames = np.ascontiguousarray(frames, np.uint8).reshape(-1, (nbits + 7) // 8)
    return np.array([lib().oc_bch_check(_p(f), nbits, framesize, t) for f in frames], np.int32)


def demap(sym, n_mod, pts, col, fec_len, qpsk_scale, psk8_scale, gain, n_sym=None):
    """sym float32 [B, L, 2] -> int16 [B, fec_len]."""
    sym = np.ascontiguousarray(sym, np.float32)
    B, L = sym.shape[0], sym.shape[1]
    n_sym = L if n_sym is None else n_sym
    out = np.zeros((B, fec_len), np.int16)
    col = np.ascontiguousarray(col, np.uint32)
    p = None if pts is None else np.ascontiguousarray(pts, np.float32)
    n_points = 0 if p is None else len(p) // 2
    for b in range(B):
        lib().oc_demap(_p(sym[b]), n_sym, n_mod, _p(p), n_points, _p(col), fec_len,
                       qpsk_scale, psk8_scale, gain, _p(out[b]))
    return out


def exp_neg(x):
    return lib().oc_exp_neg(float(x))


def log_pos(x):
    return lib().oc_log_pos(float(x))
