"""cusat_amd -- MI355X (gfx950) RX FEC hot path of cuSat.

Python access to the C ABI of ``include/cusat_fec.h`` (libcusat_fec.so,
built in-tree by ``cusat_amd/build.py``).  Device buffers are passed as raw
pointers; torch tensors are accepted wherever a device pointer is expected
(``tensor.data_ptr()``), and the stream defaults to torch's current stream.

The library must be built; importing works without it, but every call
raises if ``libcusat_fec.so`` is missing -- there is no fallback path.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcusat_fec.so")
DATA_DIR = os.path.join(_PKG, "data")

FECFRAME_NORMAL, FECFRAME_SHORT, FECFRAME_MEDIUM = 0, 1, 2
MODULATIONS = ["QPSK", "8PSK", "8APSK", "16APSK", "32APSK", "64APSK", "128APSK", "256APSK"]
RATES = ["1/4", "1/3", "2/5", "1/2", "3/5", "2/3", "3/4", "4/5", "5/6", "8/9", "9/10",
         "11/45", "4/15", "13/45", "14/45", "9/20", "7/15", "90/180", "8/15", "96/180",
         "11/20", "100/180", "26/45", "104/180", "18/30", "28/45", "23/36", "116/180",
         "20/30", "124/180", "25/36", "32/45", "128/180", "13/18", "22/30", "132/180",
         "135/180", "7/9", "140/180", "154/180"]
CTX_NO_SPLIT_KERNEL = 0x2
CTX_SEQUENTIAL_SCHEDULE = 0x1


def rate_id(r):
    return r if isinstance(r, int) else RATES.index(r)


def mod_id(m):
    return m if isinstance(m, int) else MODULATIONS.index(m.upper())


class FecInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "n_ldpc", "k_ldpc", "k_bch", "n_bch", "bch_t", "bch_degree", "q", "n_steps", "lanes",
        "cnl", "lt_alloc", "n_mod", "xfec_len", "demap_len", "max_batch", "max_trials",
        "n_collision_slots", "lds_bytes")] + [("llr_dmax", ctypes.c_float), ("llr_gain_cap", ctypes.c_float)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def lib():
    """The loaded libcusat_fec.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcusat_fec.so is not built: run python cusat_amd/build.py")
        # torch ships its own libamdhip64.so.7; load it first so this library
        # binds to the same HIP runtime instance (streams and device pointers
        # from torch are then valid here).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f, u = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint
        L.cusat_last_error.restype = ctypes.c_char_p
        L.cusat_fec_ctx_create.argtypes = [ctypes.POINTER(vp), i, i, i, i, i, u]
        L.cusat_fec_ctx_destroy.argtypes = [vp]
        L.cusat_fec_ctx_info.argtypes = [vp, ctypes.POINTER(FecInfo)]
        L.cusat_fec_query.argtypes = [i, i, i, u, ctypes.POINTER(FecInfo)]
        L.cusat_fec_demap_scales.argtypes = [vp, f, f, ctypes.POINTER(f), ctypes.POINTER(f),
                                             ctypes.POINTER(f)]
        L.cusat_fec_constellation.argtypes = [vp, vp, vp]
        L.cusat_fec_demap.argtypes = [vp, vp, vp, i, f, f, vp]
        L.cusat_ldpc_decode.argtypes = [vp, vp, vp, vp, i, vp]
        L.cusat_ldpc_decode_noalloc.argtypes = [vp, vp, vp, vp, i, vp]
        L.cusat_bch_check.argtypes = [vp, vp, vp, i, vp]
        L.cusat_rx_fec_batch.argtypes = [vp, vp, vp, i, f, f, vp, vp, vp, vp]
        L.cusat_gen_frames.argtypes = [i, i, i, ctypes.c_uint64, i, f, vp, vp, vp]
        L.cusat_modcod_constellation.argtypes = [i, i, i, vp, vp, ctypes.POINTER(i),
                                                 ctypes.POINTER(i)]
        L.cusat_bch_encode.argtypes = [i, i, vp, vp, i]
        L.cusat_ldpc_set_params.argtypes = [vp, vp, vp, i]
        L.cusat_ldpc_decode_learned.argtypes = [vp, vp, vp, vp, i, i, vp]
        L.cusat_rx_fec_learned.argtypes = [vp, vp, vp, vp, vp, vp, i, i, vp]
        L.cusat_learn_ctx_create.argtypes = [ctypes.POINTER(vp), vp, i, i]
        L.cusat_learn_ctx_destroy.argtypes = [vp]
        L.cusat_learn_fwd_bwd.argtypes = [vp, vp, vp, i, vp, vp, f, u, vp, vp, vp, vp, vp, vp]
        L.cusat_learn_adam.argtypes = [vp, vp, vp, vp, i, f, f, f, f, i, f, f, vp]
        L.cusat_bb_extract.argtypes = [vp, vp, vp, vp, vp, vp, i, vp]
        L.cusat_bb_prbs.argtypes = [vp, i]
        L.cusat_bb_crc8.argtypes = [vp, i]
        L.cusat_encode_frames.argtypes = [i, i, i, ctypes.c_uint64, i, f, vp, vp, vp]
        L.cusat_gen_frames_dev.argtypes = [vp, ctypes.c_uint64, i, f, vp, vp, vp, vp, vp]
        L.cusat_stream_create.argtypes = [ctypes.POINTER(vp), vp, i, i, u]
        L.cusat_stream_destroy.argtypes = [vp]
        L.cusat_stream_plan.argtypes = [vp, vp, i, vp]
        L.cusat_stream_plan_async.argtypes = [vp, vp, i, vp]
        L.cusat_stream_rx_plan_async.argtypes = [vp, vp, i, vp, i, vp]
        L.cusat_stream_rx_dd.argtypes = [vp, vp, vp, vp, vp]
        L.cusat_stream_rx_dd_estimate.argtypes = [vp, vp, vp]
        L.cusat_stream_rx_dd_resolve.argtypes = [vp, vp, vp, vp]
        L.cusat_stream_rx.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.cusat_stream_rx_demap.argtypes = [vp, vp, vp]
        L.cusat_stream_rx_decode.argtypes = [vp, vp, vp, vp, vp]
        L.cusat_stream_rx_bb.argtypes = [vp, vp, vp, vp, vp, vp]
        L.cusat_stream_rx_work.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.cusat_stream_get_tuning.argtypes = [vp, ctypes.POINTER(StreamTuning), ctypes.POINTER(i)]
        L.cusat_stream_decode.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.cusat_stream_autotune.argtypes = [vp, vp, vp, vp, vp, vp, i, vp, ctypes.POINTER(StreamTuning),
                                            ctypes.POINTER(f)]
        _lib = L
    return _lib


def copy_from_device(addr, nbytes):
    """Host copy (uint8 array) of nbytes of device memory at the integer
    address addr, e.g. a buffer of StreamContext.rx_work(); synchronous
    hipMemcpy through the HIP runtime this library is linked to."""
    L = lib()
    fn = L.hipMemcpy
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(int(nbytes), np.uint8)
    if nbytes:
        rc = fn(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(int(addr)), int(nbytes), 4)  # hipMemcpyDefault
        if rc != 0:
            raise RuntimeError("hipMemcpy from device failed (%d)" % rc)
    return out


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, lib().cusat_last_error().decode()))


def _ptr(x):
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr())
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(int(x))


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


class FecContext:
    """One MODCOD's decoder state on the current HIP device
    (demapper_cuda + ldpc_decode_cuda + bch_check_cuda of the reference)."""

    def __init__(self, framesize, rate, modulation, max_trials=50, max_batch=64, flags=0):
        self.framesize, self.rate, self.modulation = framesize, rate_id(rate), mod_id(modulation)
        h = ctypes.c_void_p()
        _check(lib().cusat_fec_ctx_create(ctypes.byref(h), framesize, self.rate, self.modulation,
                                          max_trials, max_batch, flags), "cusat_fec_ctx_create")
        self._destroy = lib().cusat_fec_ctx_destroy
        self._h = h
        inf = FecInfo()
        _check(lib().cusat_fec_ctx_info(h, ctypes.byref(inf)), "cusat_fec_ctx_info")
        self.info = inf.as_dict()

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            # the bound function, not lib(): at interpreter exit the module's
            # globals may already be gone when the destructor runs
            self._destroy(h)

    __del__ = close

    def demap_scales(self, noise_var, llr_gain):
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        _check(lib().cusat_fec_demap_scales(self._h, noise_var, llr_gain, ctypes.byref(a),
                                            ctypes.byref(b), ctypes.byref(c)), "demap_scales")
        return a.value, b.value, c.value

    def constellation(self):
        pts = np.zeros(2 << self.info["n_mod"], np.float32)
        col = np.zeros(8, np.uint32)
        _check(lib().cusat_fec_constellation(self._h, _ptr(pts), _ptr(col)), "constellation")
        return pts, col

    def demap(self, sym, llr, batch, noise_var, llr_gain, stream=None):
        _check(lib().cusat_fec_demap(self._h, _ptr(sym), _ptr(llr), batch, noise_var, llr_gain,
                                     _stream(stream)), "cusat_fec_demap")

    def decode(self, llr, hard, trials, batch, stream=None):
        _check(lib().cusat_ldpc_decode(self._h, _ptr(llr), _ptr(hard), _ptr(trials), batch,
                                       _stream(stream)), "cusat_ldpc_decode")

    def bch_check(self, hard, valid, batch, stream=None):
        _check(lib().cusat_bch_check(self._h, _ptr(hard), _ptr(valid), batch, _stream(stream)),
               "cusat_bch_check")

    def rx_fec_batch(self, sym, llr, batch, noise_var, llr_gain, hard, trials, valid, stream=None):
        _check(lib().cusat_rx_fec_batch(self._h, _ptr(sym), _ptr(llr), batch, noise_var, llr_gain,
                                        _ptr(hard), _ptr(trials), _ptr(valid), _stream(stream)),
               "cusat_rx_fec_batch")

    def set_params(self, beta, alpha):
        """Learned check-node parameters: int16 [n_iters, q] offsets (LLR
        units) and scales (n/64, 0 = none); empty arrays clear them."""
        beta = np.ascontiguousarray(beta, np.int16)
        alpha = np.ascontiguousarray(alpha, np.int16)
        n = 0 if beta.size == 0 else beta.reshape(-1, self.info["q"]).shape[0]
        assert alpha.shape == beta.shape
        _check(lib().cusat_ldpc_set_params(self._h, _ptr(beta) if n else None,
                                           _ptr(alpha) if n else None, n), "cusat_ldpc_set_params")

    def decode_learned(self, llr, hard, trials, batch, max_trials, stream=None):
        _check(lib().cusat_ldpc_decode_learned(self._h, _ptr(llr), _ptr(hard), _ptr(trials), batch,
                                               max_trials, _stream(stream)),
               "cusat_ldpc_decode_learned")

    def gen_frames_dev(self, seed, batch, es_n0_db, sym_out, info_in=None, info_out=None, cw_out=None,
                       stream=None):
        """Test traffic on the GPU (cusat_gen_frames_dev): device tensors
        sym_out [B, xfec, 2] float32, optional info_in / info_out
        [B, k_bch/8] uint8 and cw_out [B, n_ldpc/8] uint8."""
        _check(lib().cusat_gen_frames_dev(self._h, int(seed) & 0xffffffffffffffff, batch, float(es_n0_db),
                                          _ptr(info_in), _ptr(sym_out), _ptr(info_out), _ptr(cw_out),
                                          _stream(stream)), "cusat_gen_frames_dev")

    def bb_extract(self, hard, trials, valid, out, dfl, batch, stream=None):
        """Post-BCH BB descramble + BBHEADER check + GSE data field (GPU)."""
        _check(lib().cusat_bb_extract(self._h, _ptr(hard), _ptr(trials), _ptr(valid), _ptr(out), _ptr(dfl),
                                      batch, _stream(stream)), "cusat_bb_extract")

    def rx_fec_learned(self, llr, hard, trials, valid, fallback, batch, learned_trials, stream=None):
        _check(lib().cusat_rx_fec_learned(self._h, _ptr(llr), _ptr(hard), _ptr(trials), _ptr(valid),
                                          _ptr(fallback), batch, learned_trials, _stream(stream)),
               "cusat_rx_fec_learned")


LEARN_QUANTIZED = 0x1
STREAM_LEARNED = 0x1
STREAM_NO_PAIRS = 0x2
STREAM_ORDER_GIVEN, STREAM_ORDER_COST = 0, 1
STREAM_FRAME = np.dtype([("code", np.int32), ("reserved", np.int32), ("llr_offset", np.int64),
                         ("hard_offset", np.int64)])
# cusat_stream_rx_frame: a frame of a mixed receive batch
STREAM_RX_FRAME = np.dtype([("code", np.int32), ("flags", np.int32), ("sym_offset", np.int64),
                            ("out_offset", np.int64), ("noise_var", np.float32), ("llr_gain", np.float32)])
RX_SCALE_STATE = 1
# cusat_stream_rx_event: an SNR event of a receive batch
STREAM_RX_EVENT = np.dtype([("code", np.int32), ("frame", np.int32), ("snr_lin", np.float32),
                            ("snr_db", np.float32)])


class StreamTuning(ctypes.Structure):
    _fields_ = [("wg_per_cu", ctypes.c_int), ("frames_per_job", ctypes.c_int), ("order", ctypes.c_int),
                ("split_edges", ctypes.c_int)]

    def as_dict(self):
        return {"wg_per_cu": self.wg_per_cu, "frames_per_job": self.frames_per_job, "order": self.order,
                "split_edges": self.split_edges}


class StreamContext:
    """Persistent mixed-stream decoder over several FecContexts
    (cusat_stream_* of the C ABI).  frames: STREAM_FRAME records."""

    def __init__(self, contexts, max_frames, flags=0):
        self.contexts = list(contexts)  # keep them alive
        arr = (ctypes.c_void_p * len(self.contexts))(*[c._h.value for c in self.contexts])
        h = ctypes.c_void_p()
        _check(lib().cusat_stream_create(ctypes.byref(h), ctypes.cast(arr, ctypes.c_void_p), len(arr),
                                         max_frames, flags), "cusat_stream_create")
        self._destroy = lib().cusat_stream_destroy
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            # the bound function, not lib(): at interpreter exit the module's
            # globals may already be gone when the destructor runs
            self._destroy(h)

    __del__ = close

    def plan(self, frames, tuning=None):
        frames = np.ascontiguousarray(frames, STREAM_FRAME)
        t = None
        if tuning is not None:
            t = StreamTuning(tuning["wg_per_cu"], tuning["frames_per_job"], tuning["order"],
                             tuning.get("split_edges", 0))
        _check(lib().cusat_stream_plan(self._h, _ptr(frames), len(frames),
                                       ctypes.byref(t) if t is not None else None), "cusat_stream_plan")

    def plan_async(self, frames, stream=None):
        """New frames under the current tuning, copied on the stream (no host wait)."""
        frames = np.ascontiguousarray(frames, STREAM_FRAME)
        _check(lib().cusat_stream_plan_async(self._h, _ptr(frames), len(frames), _stream(stream)),
               "cusat_stream_plan_async")

    def rx_plan_async(self, frames, events=None, stream=None):
        """Make a mixed receive batch (STREAM_RX_FRAME records, optional
        STREAM_RX_EVENT SNR events) current."""
        frames = np.ascontiguousarray(frames, STREAM_RX_FRAME)
        ev = np.ascontiguousarray(events if events is not None else np.zeros(0, STREAM_RX_EVENT), STREAM_RX_EVENT)
        _check(lib().cusat_stream_rx_plan_async(self._h, _ptr(frames), len(frames), _ptr(ev), len(ev),
                                                _stream(stream)), "cusat_stream_rx_plan_async")

    def rx_dd(self, sym, state, state_out=None, stream=None):
        """Apply the batch's SNR events to state (float32 [n_codes, 4], device)."""
        _check(lib().cusat_stream_rx_dd(self._h, _ptr(sym), _ptr(state), _ptr(state_out), _stream(stream)),
               "cusat_stream_rx_dd")

    def rx_dd_estimate(self, sym, stream=None):
        """The estimate launch of rx_dd (reads the symbols and the plan only)."""
        _check(lib().cusat_stream_rx_dd_estimate(self._h, _ptr(sym), _stream(stream)), "cusat_stream_rx_dd_estimate")

    def rx_dd_resolve(self, state, state_out=None, stream=None):
        """The resolve launch of rx_dd: the events applied in order to state."""
        _check(lib().cusat_stream_rx_dd_resolve(self._h, _ptr(state), _ptr(state_out), _stream(stream)),
               "cusat_stream_rx_dd_resolve")

    def rx(self, sym, out, trials, valid, dfl, stream=None):
        """Demap, decode, BCH-check and BB-extract the current receive batch."""
        _check(lib().cusat_stream_rx(self._h, _ptr(sym), _ptr(out), _ptr(trials), _ptr(valid), _ptr(dfl),
                                     _stream(stream)), "cusat_stream_rx")

    def rx_demap(self, sym, stream=None):
        """First launch of rx(): the mixed demap into the LLR work buffer."""
        _check(lib().cusat_stream_rx_demap(self._h, _ptr(sym), _stream(stream)), "cusat_stream_rx_demap")

    def rx_decode(self, trials, valid, fallback=None, stream=None):
        """Second launch of rx(): persistent decode + BCH of the LLR work buffer."""
        _check(lib().cusat_stream_rx_decode(self._h, _ptr(trials), _ptr(valid), _ptr(fallback), _stream(stream)),
               "cusat_stream_rx_decode")

    def rx_bb(self, out, trials, valid, dfl, stream=None):
        """Third launch of rx(): BB descramble + data-field extraction."""
        _check(lib().cusat_stream_rx_bb(self._h, _ptr(out), _ptr(trials), _ptr(valid), _ptr(dfl),
                                        _stream(stream)), "cusat_stream_rx_bb")

    def rx_work(self):
        """Device addresses (ints) of the rx LLR and hard-decision work buffers."""
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().cusat_stream_rx_work(self._h, ctypes.byref(a), ctypes.byref(b)), "cusat_stream_rx_work")
        return a.value, b.value

    def tuning(self):
        t, m = StreamTuning(), ctypes.c_int()
        _check(lib().cusat_stream_get_tuning(self._h, ctypes.byref(t), ctypes.byref(m)),
               "cusat_stream_get_tuning")
        d = t.as_dict()
        d["max_wg_per_cu"] = m.value
        return d

    def decode(self, llr, hard, trials, valid, fallback=None, stream=None):
        _check(lib().cusat_stream_decode(self._h, _ptr(llr), _ptr(hard), _ptr(trials), _ptr(valid),
                                         _ptr(fallback), _stream(stream)), "cusat_stream_decode")

    def autotune(self, llr, hard, trials, valid, fallback=None, reps=3, stream=None):
        t, ms = StreamTuning(), ctypes.c_float()
        _check(lib().cusat_stream_autotune(self._h, _ptr(llr), _ptr(hard), _ptr(trials), _ptr(valid),
                                           _ptr(fallback), reps, _stream(stream), ctypes.byref(t),
                                           ctypes.byref(ms)), "cusat_stream_autotune")
        return t.as_dict(), ms.value


class LearnContext:
    """fp32 forward/backward state for training the learned parameters of
    one code over n_iters unrolled iterations (cusat_learn_* of the C ABI)."""

    def __init__(self, fec, max_batch, n_iters):
        self.fec, self.max_batch, self.n_iters = fec, max_batch, n_iters
        h = ctypes.c_void_p()
        _check(lib().cusat_learn_ctx_create(ctypes.byref(h), fec._h, max_batch, n_iters),
               "cusat_learn_ctx_create")
        self._destroy = lib().cusat_learn_ctx_destroy
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            # the bound function, not lib(): at interpreter exit the module's
            # globals may already be gone when the destructor runs
            self._destroy(h)

    __del__ = close

    def fwd_bwd(self, llr, cw, batch, alpha, beta, kappa, grad_alpha, grad_beta, loss,
                app_out=None, grad_llr=None, quantized=False, stream=None):
        _check(lib().cusat_learn_fwd_bwd(self._h, _ptr(llr), _ptr(cw), batch, _ptr(alpha), _ptr(beta),
                                         kappa, LEARN_QUANTIZED if quantized else 0,
                                         _ptr(grad_alpha), _ptr(grad_beta), _ptr(loss),
                                         _ptr(app_out), _ptr(grad_llr), _stream(stream)),
               "cusat_learn_fwd_bwd")


def adam_step(param, grad, m, v, lr, step, lo, hi, b1=0.9, b2=0.999, eps=1e-8, stream=None):
    """In-place Adam update of a device float tensor (cusat_learn_adam)."""
    _check(lib().cusat_learn_adam(_ptr(param), _ptr(grad), _ptr(m), _ptr(v), param.numel(), lr, b1, b2,
                                  eps, step, lo, hi, _stream(stream)), "cusat_learn_adam")


def gen_frames(framesize, rate, modulation, seed, batch, es_n0_db, want_cw=False):
    """Seeded transmit chain on the host CPU.  Returns (symbols float32
    [B, xfec, 2], info bytes uint8 [B, k_bch/8], codeword bytes or None)."""
    rate, modulation = rate_id(rate), mod_id(modulation)
    info = geometry(framesize, rate, modulation)
    sym = np.zeros((batch, info["xfec_len"], 2), np.float32)
    inf = np.zeros((batch, info["k_bch"] // 8), np.uint8)
    cw = np.zeros((batch, info["n_ldpc"] // 8), np.uint8) if want_cw else None
    _check(lib().cusat_gen_frames(framesize, rate, modulation, seed, batch, es_n0_db, _ptr(sym),
                                  _ptr(inf), _ptr(cw)), "cusat_gen_frames")
    return sym, inf, cw


def geometry(framesize, rate, modulation, flags=0):
    """Code/modulation dimensions from the host tables (no device needed)."""
    inf = FecInfo()
    _check(lib().cusat_fec_query(framesize, rate_id(rate), mod_id(modulation), flags,
                                 ctypes.byref(inf)), "cusat_fec_query")
    return inf.as_dict()


def modcod_constellation(framesize, rate, modulation):
    """Host tables of a MODCOD: (points float32 [2 << n_mod], column offsets
    uint32 [8], bits per symbol, interleaver rows)."""
    pts = np.zeros(512, np.float32)
    col = np.zeros(8, np.uint32)
    nm, rows = ctypes.c_int(), ctypes.c_int()
    _check(lib().cusat_modcod_constellation(framesize, rate_id(rate), mod_id(modulation), _ptr(pts),
                                            _ptr(col), ctypes.byref(nm), ctypes.byref(rows)),
           "cusat_modcod_constellation")
    return pts[: 2 << nm.value].copy(), col, nm.value, rows.value


def bch_encode(framesize, rate, msg):
    """Systematic BCH encode of uint8 [B, k_bch/8] messages (host)."""
    g = geometry(framesize, rate, 0)
    msg = np.ascontiguousarray(msg, np.uint8).reshape(-1, g["k_bch"] // 8)
    cw = np.zeros((msg.shape[0], g["n_bch"] // 8), np.uint8)
    _check(lib().cusat_bch_encode(framesize, rate_id(rate), _ptr(msg), _ptr(cw), msg.shape[0]),
           "cusat_bch_encode")
    return cw


def frame_passes_fec_checks(trials, bch_valid):
    """Frame acceptance of rx_fec_chain::frame_passes_fec_checks
    (lib/cusat_dsp/rx_fec_chain.h:193-196): the LDPC decoder converged
    (trials >= 0) and the BCH syndrome is zero.  Element-wise on arrays."""
    return (np.asarray(trials) >= 0) & (np.asarray(bch_valid) != 0)


def bb_prbs(n_bytes):
    """The BBFRAME scrambling sequence (EN 302 307-1 clause 5.2.2), bytes MSB first."""
    out = np.zeros(n_bytes, np.uint8)
    _check(lib().cusat_bb_prbs(_ptr(out), n_bytes), "cusat_bb_prbs")
    return out


def bb_crc8(data):
    d = np.ascontiguousarray(data, np.uint8)
    r = lib().cusat_bb_crc8(_ptr(d), len(d))
    if r < 0:
        raise RuntimeError("cusat_bb_crc8 failed")
    return r


def encode_frames(framesize, rate, modulation, seed, info, es_n0_db, want_cw=False):
    """Transmit chain from given BBFRAMEs uint8 [B, k_bch/8] (host)."""
    g = geometry(framesize, rate, modulation)
    info = np.ascontiguousarray(info, np.uint8).reshape(-1, g["k_bch"] // 8)
    B = info.shape[0]
    sym = np.zeros((B, g["xfec_len"], 2), np.float32)
    cw = np.zeros((B, g["n_ldpc"] // 8), np.uint8) if want_cw else None
    _check(lib().cusat_encode_frames(framesize, rate_id(rate), mod_id(modulation), seed, B, es_n0_db,
                                     _ptr(info), _ptr(sym), _ptr(cw)), "cusat_encode_frames")
    return sym, cw
