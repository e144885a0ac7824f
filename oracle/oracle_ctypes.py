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
