"""ctypes binding of the C ABI in ``include/fbfuse_hip.h``.

This is how Python (tests, bench, smoke) reaches the native library.  The
library is loaded from this package directory; if it is missing the import of
any compute entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfbfuse_hip.so")

FBF_OK = 0
FBF_ERR_ARG = 1
FBF_ERR_SIZE = 2
FBF_ERR_HIP = 3
FBF_ERR_UNSUPPORTED = 4
FBF_ERR_STATE = 5
FBF_ERR_CAPACITY = 6

QUANT_ROUND = 0
QUANT_TRUNC = 1


class FbfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[fbf status {code}] {msg}")
        self.code = code


class CbDesc(C.Structure):
    _fields_ = [("nantennas", C.c_int), ("nbeams", C.c_int), ("nchans", C.c_int),
                ("tscrunch", C.c_int), ("fscrunch", C.c_int),
                ("nsamples_per_heap_out", C.c_int), ("quant_mode", C.c_int)]


class SubarrayDesc(C.Structure):
    _fields_ = [("start_antenna", C.c_int), ("nantennas", C.c_int), ("nbeams", C.c_int),
                ("tscrunch", C.c_int), ("fscrunch", C.c_int), ("nsamples_per_heap_out", C.c_int),
                ("quant_mode", C.c_int)]


class FeedDesc(C.Structure):
    _fields_ = [("nantennas", C.c_int), ("nbeams", C.c_int), ("nchans", C.c_int), ("nheap_groups", C.c_int),
                ("nsamples_per_heap", C.c_int), ("tscrunch", C.c_int), ("fscrunch", C.c_int),
                ("nsamples_per_heap_out", C.c_int), ("quant_mode", C.c_int), ("ib_tscrunch", C.c_int),
                ("ib_fscrunch", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("in_flight", C.c_uint64), ("pending", C.c_uint64), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mean_ms", C.c_double),
                ("input_bytes", C.c_double), ("wall_s", C.c_double), ("hbm_budget", C.c_size_t),
                ("hbm_peak", C.c_size_t)]


class IbDesc(C.Structure):
    _fields_ = [("nantennas", C.c_int), ("nchans", C.c_int), ("nsamples_per_heap", C.c_int),
                ("tscrunch", C.c_int), ("fscrunch", C.c_int), ("quant_mode", C.c_int)]


vp = C.c_void_p
sz = C.c_size_t
i32 = C.c_int
pvp = C.POINTER(C.c_void_p)

# name -> argtypes; every function returns int except fbf_last_error.
_SIGS = {
    "fbf_version": [],
    "fbf_device_count": [C.POINTER(C.c_int)],
    "fbf_set_device": [i32],
    "fbf_device_name": [C.c_char_p, sz],
    "fbf_malloc": [pvp, sz],
    "fbf_free": [vp],
    "fbf_host_malloc": [pvp, sz],
    "fbf_host_free": [vp],
    "fbf_memcpy_h2d": [vp, vp, sz, vp],
    "fbf_memcpy_d2h": [vp, vp, sz, vp],
    "fbf_memcpy_d2d": [vp, vp, sz, vp],
    "fbf_memset": [vp, i32, sz, vp],
    "fbf_stream_create": [pvp],
    "fbf_stream_destroy": [vp],
    "fbf_stream_synchronize": [vp],
    "fbf_device_synchronize": [],
    "fbf_event_create": [pvp],
    "fbf_event_destroy": [vp],
    "fbf_event_record": [vp, vp],
    "fbf_stream_wait_event": [vp, vp],
    "fbf_event_synchronize": [vp],
    "fbf_event_elapsed_ms": [vp, vp, C.POINTER(C.c_float)],
    "fbf_synth_gaussian_int8": [C.c_uint64, C.c_double, i32, vp, sz],
    "fbf_synth_uniform_f32": [C.c_uint64, C.c_float, C.c_float, vp, sz],
    "fbf_power_levels": [C.c_float, C.c_float, i32, i32, i32, i32, i32, i32, i32,
                         C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "fbf_channel_frequencies": [C.c_float, C.c_float, i32, vp],
    "fbf_channel_frequencies_f64": [C.c_double, C.c_double, i32, vp],
    "fbf_split_transpose": [vp, sz, vp, sz, i32, i32, i32, i32, i32, vp],
    "fbf_cb_plan_create": [C.POINTER(CbDesc), vp, vp, pvp],
    "fbf_cb_plan_set_levels": [vp, vp, vp],
    "fbf_cb_plan_destroy": [vp],
    "fbf_cb_plan_set_variant": [vp, i32],
    "fbf_cb_variant_count": [C.POINTER(C.c_int)],
    "fbf_cb_output_bytes": [vp, sz, C.POINTER(C.c_size_t)],
    "fbf_cb_beamform": [vp, vp, sz, vp, sz, vp, sz, vp],
    "fbf_cb_beam_sums": [vp, vp, sz, vp, sz, vp, sz, vp],
    "fbf_cb_beam_power": [vp, vp, sz, vp, sz, vp, sz, vp],
    "fbf_cb_beamform_subarray": [vp, vp, sz, i32, i32, vp, sz, vp, sz, vp],
    "fbf_cb_beamform_taftp": [vp, vp, sz, i32, i32, i32, vp, sz, vp, sz, vp],
    "fbf_cb_taftp_direct": [vp, i32, i32, C.c_longlong, C.POINTER(C.c_int)],
    "fbf_cb_ib_beamform_taftp": [vp, vp, vp, sz, i32, vp, sz, vp, sz, vp, sz, vp],
    "fbf_cb_ib_fused": [vp, vp, i32, C.c_longlong, C.POINTER(C.c_int)],
    "fbf_cb_ib_beamform_taftp_next": [vp, vp, vp, sz, i32, vp, sz, vp, sz, vp, sz, vp, vp, sz, C.c_double, vp, sz,
                                      vp],
    "fbf_cb_plan_set_chunk": [vp, i32],
    "fbf_cb_plan_set_timing": [vp, vp, vp],
    "fbf_cb_autotune": [vp, C.c_longlong, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)],
    "fbf_batch_create": [i32, i32, i32, i32, i32, C.POINTER(SubarrayDesc), C.POINTER(C.c_void_p),
                         C.POINTER(C.c_void_p), pvp],
    "fbf_batch_destroy": [vp],
    "fbf_batch_enable_ib": [vp, i32, i32, i32, vp, vp],
    "fbf_batch_output_bytes": [vp, i32, C.POINTER(C.c_size_t)],
    "fbf_batch_weights_bytes": [vp, i32, C.POINTER(C.c_size_t)],
    "fbf_batch_set_variant": [vp, i32, i32, i32],
    "fbf_batch_autotune": [vp, i32],
    "fbf_batch_set_serial": [vp, i32],
    "fbf_batch_get_serial": [vp, vp],
    "fbf_batch_run": [vp, vp, sz, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), i32, vp],
    "fbf_cb_beamform_ws": [vp, vp, sz, vp, sz, vp, sz, vp, sz, vp],
    "fbf_cb_beamform_noalloc": [vp, vp, sz, vp, sz, vp, sz, vp],
    "fbf_split_transpose_ib": [vp, sz, vp, sz, i32, i32, vp, vp, sz, vp],
    "fbf_split_transpose_ib_weights": [vp, sz, vp, sz, i32, i32, vp, vp, sz, vp, vp, sz, C.c_double, vp, sz, vp],
    "fbf_tuse_transpose": [vp, sz, vp, sz, i32, i32, i32, i32, i32, i32, i32, vp],
    "fbf_feng_to_dada": [vp, vp, sz, i32, vp],
    "fbf_cb_plan_reserve": [vp, C.c_longlong],
    "fbf_server_create": [i32, sz, pvp],
    "fbf_server_destroy": [vp],
    "fbf_server_add_feed": [vp, C.POINTER(FeedDesc), vp, vp, vp, vp, C.POINTER(C.c_int)],
    "fbf_server_feed_bytes": [vp, i32, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "fbf_server_set_weights": [vp, i32, vp, sz],
    "fbf_server_set_delays": [vp, i32, vp, sz, vp, C.c_double],
    "fbf_server_start": [vp],
    "fbf_server_submit": [vp, i32, vp, sz, i32, vp, vp, C.c_uint64],
    "fbf_server_poll": [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_double), i32],
    "fbf_server_stats_get": [vp, C.POINTER(ServerStats)],
    "fbf_server_reset_stats": [vp],
    "fbf_server_drain": [vp],
    "fbf_ib_plan_create": [C.POINTER(IbDesc), vp, vp, pvp],
    "fbf_ib_plan_set_levels": [vp, vp, vp],
    "fbf_ib_plan_destroy": [vp],
    "fbf_ib_output_bytes": [vp, sz, C.POINTER(C.c_size_t)],
    "fbf_ib_beamform": [vp, vp, sz, vp, sz, vp],
    "fbf_ib_beamform_weights": [vp, vp, sz, vp, sz, vp, vp, sz, C.c_double, vp, sz, vp],
    "fbf_ib_power": [vp, vp, sz, vp, sz, vp],
    "fbf_weights_plan_create": [i32, i32, i32, vp, pvp],
    "fbf_weights_plan_destroy": [vp],
    "fbf_weights_generate": [vp, vp, sz, C.c_double, vp, sz, vp, vp],
    "fbf_weights_generate_steps": [vp, vp, sz, C.c_double, C.c_double, i32, vp, sz, vp],
}

_lib = None


def header_symbols():
    """Names of every function declared in include/*.h (parsed from the text)."""
    import re
    inc = os.path.join(os.path.dirname(PKG), "include")
    names = []
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            with open(os.path.join(inc, fn)) as f:
                txt = f.read()
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(fbf_[a-z0-9_]+)\s*\(", txt, re.M)
    return names


def lib():
    """Load libfbfuse_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"native library {LIB_PATH} is missing: run __graft_entry__.build(); "
                           "there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    L.fbf_last_error.restype = C.c_char_p
    L.fbf_last_error.argtypes = []
    for name, args in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = C.c_int
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status != FBF_OK:
        msg = lib().fbf_last_error()
        raise FbfError(status, msg.decode() if msg else "")
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# Small RAII helpers
# ---------------------------------------------------------------------------
class Stream:
    def __init__(self):
        h = C.c_void_p()
        call("fbf_stream_create", C.byref(h))
        self.handle = h

    def sync(self):
        call("fbf_stream_synchronize", self.handle)

    def wait(self, event):
        """Later work on this stream waits for `event` (fbf_stream_wait_event)."""
        call("fbf_stream_wait_event", self.handle, event.handle)

    def close(self):
        if self.handle:
            call("fbf_stream_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        call("fbf_event_create", C.byref(h))
        self.handle = h

    def record(self, stream):
        call("fbf_event_record", self.handle, stream.handle if stream else None)

    def sync(self):
        call("fbf_event_synchronize", self.handle)

    def elapsed_ms(self, later):
        ms = C.c_float()
        call("fbf_event_elapsed_ms", self.handle, later.handle, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                call("fbf_event_destroy", self.handle)
        except Exception:
            pass


class DeviceBuffer:
    """Device allocation of `nbytes` bytes (hipMalloc through the C ABI)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        call("fbf_malloc", C.byref(h), self.nbytes)
        self.ptr = h

    @classmethod
    def from_array(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr, stream)
        return b

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        s = stream.handle if stream else None
        call("fbf_memcpy_h2d", self.ptr, _np_ptr(arr), arr.nbytes, s)
        if stream:
            stream.sync()
        else:
            call("fbf_device_synchronize")

    def download(self, dtype=np.int8, count=None, stream=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        s = stream.handle if stream else None
        call("fbf_memcpy_d2h", _np_ptr(out), self.ptr, out.nbytes, s)
        if stream:
            stream.sync()
        else:
            call("fbf_device_synchronize")
        return out

    def fill(self, value=0, stream=None):
        s = stream.handle if stream else None
        call("fbf_memset", self.ptr, int(value), self.nbytes, s)

    def free(self):
        if self.ptr:
            call("fbf_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Host helpers that call native code
# ---------------------------------------------------------------------------
def synth_gaussian_int8(seed, sigma, n, clip=True):
    out = np.empty(int(n), dtype=np.int8)
    call("fbf_synth_gaussian_int8", C.c_uint64(seed), float(sigma), int(bool(clip)), _np_ptr(out), out.size)
    return out


def synth_uniform_f32(seed, lo, hi, n):
    out = np.empty(int(n), dtype=np.float32)
    call("fbf_synth_uniform_f32", C.c_uint64(seed), float(lo), float(hi), _np_ptr(out), out.size)
    return out


def power_levels(input_level, output_level, cb_nantennas, cb_tscrunch, cb_fscrunch,
                 ib_nantennas, ib_tscrunch, ib_fscrunch, npol=2):
    v = [C.c_float() for _ in range(4)]
    call("fbf_power_levels", input_level, output_level, cb_nantennas, cb_tscrunch, cb_fscrunch,
         ib_nantennas, ib_tscrunch, ib_fscrunch, npol, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def channel_frequencies(cfreq, bw, nchans):
    out = np.empty(nchans, dtype=np.float32)
    call("fbf_channel_frequencies", cfreq, bw, nchans, _np_ptr(out))
    return out


def channel_frequencies_f64(cfreq, bw, nchans):
    out = np.empty(nchans, dtype=np.float64)
    call("fbf_channel_frequencies_f64", cfreq, bw, nchans, _np_ptr(out))
    return out


def _s(stream):
    return stream.handle if stream else None


def split_transpose(taftp, ftpa, total_nantennas, used_nantennas, start_antenna, nchans,
                    nsamples_per_heap=256, stream=None):
    call("fbf_split_transpose", taftp.ptr, taftp.nbytes, ftpa.ptr, ftpa.nbytes, total_nantennas,
         used_nantennas, start_antenna, nchans, nsamples_per_heap, _s(stream))


def split_transpose_ib(taftp, ftpa, nchans, ib_plan, tf, nsamples_per_heap=256, stream=None):
    """Full-antenna split-transpose and incoherent beam in one pass."""
    call("fbf_split_transpose_ib", taftp.ptr, taftp.nbytes, ftpa.ptr, ftpa.nbytes, nchans, nsamples_per_heap,
         ib_plan.handle, tf.ptr, tf.nbytes, _s(stream))


def split_transpose_ib_weights(taftp, ftpa, nchans, ib_plan, tf, weights_plan, delays, epoch, weights,
                               nsamples_per_heap=256, stream=None):
    """split_transpose_ib and the weights generator in one launch."""
    call("fbf_split_transpose_ib_weights", taftp.ptr, taftp.nbytes, ftpa.ptr, ftpa.nbytes, nchans,
         nsamples_per_heap, ib_plan.handle, tf.ptr, tf.nbytes, weights_plan.handle, delays.ptr, delays.nbytes,
         float(epoch), weights.ptr, weights.nbytes, _s(stream))


def tuse_transpose(block, out, nchans, nsamples, nfreq, nbeams, ngroups, beam0=0, nbeams_out=None, stream=None):
    """TUSE beam corner turn of the beams [beam0, beam0 + nbeams_out):
    [ngroups][nbeams][nfreq][nsamples][nchans] -> per beam
    [ngroups][nsamples][nfreq][nchans] (device buffers)."""
    if nbeams_out is None:
        nbeams_out = nbeams - beam0
    call("fbf_tuse_transpose", block.ptr, block.nbytes, out.ptr, out.nbytes, nchans, nsamples, nfreq, nbeams,
         ngroups, beam0, nbeams_out, _s(stream))


def feng_to_dada(block, out, nchans, stream=None):
    """F-engine heap reorder, per timestamp [nchans][256] -> [256][nchans]
    32-bit words (device buffers of equal size)."""
    call("fbf_feng_to_dada", block.ptr, out.ptr, block.nbytes, nchans, _s(stream))


class CbPlan:
    """Coherent beamformer plan (fbf_cb_plan)."""

    def __init__(self, nantennas, nbeams, nchans, tscrunch, fscrunch, offsets, scales,
                 nsamples_per_heap_out=0, quant_mode=QUANT_ROUND):
        self.desc = CbDesc(nantennas, nbeams, nchans, tscrunch, fscrunch, nsamples_per_heap_out, quant_mode)
        off = np.ascontiguousarray(offsets, dtype=np.float32)
        sc = np.ascontiguousarray(scales, dtype=np.float32)
        h = C.c_void_p()
        call("fbf_cb_plan_create", C.byref(self.desc), _np_ptr(off), _np_ptr(sc), C.byref(h))
        self.handle = h

    def set_variant(self, v):
        call("fbf_cb_plan_set_variant", self.handle, int(v))

    def set_levels(self, offsets, scales):
        off = np.ascontiguousarray(offsets, dtype=np.float32)
        sc = np.ascontiguousarray(scales, dtype=np.float32)
        call("fbf_cb_plan_set_levels", self.handle, _np_ptr(off), _np_ptr(sc))

    def output_bytes(self, ftpa_bytes):
        n = C.c_size_t()
        call("fbf_cb_output_bytes", self.handle, ftpa_bytes, C.byref(n))
        return n.value

    def beamform(self, ftpa, weights, out, stream=None):
        call("fbf_cb_beamform", self.handle, ftpa.ptr, ftpa.nbytes, weights.ptr, weights.nbytes,
             out.ptr, out.nbytes, _s(stream))

    def set_timing(self, before=None, after=None):
        """Events recorded around the coherent kernel alone in later calls
        (None, None: stop)."""
        call("fbf_cb_plan_set_timing", self.handle, before.handle if before else None,
             after.handle if after else None)

    def set_chunk(self, c):
        call("fbf_cb_plan_set_chunk", self.handle, int(c))

    def autotune(self, nsamples, reps=5):
        v, c, ms = C.c_int(), C.c_int(), C.c_float()
        call("fbf_cb_autotune", self.handle, int(nsamples), int(reps), C.byref(v), C.byref(c), C.byref(ms))
        return v.value, c.value, ms.value

    def beamform_subarray(self, ftpa_block, block_nantennas, antenna_offset, weights, out, stream=None):
        call("fbf_cb_beamform_subarray", self.handle, ftpa_block.ptr, ftpa_block.nbytes, int(block_nantennas),
             int(antenna_offset), weights.ptr, weights.nbytes, out.ptr, out.nbytes, _s(stream))

    def beamform_taftp(self, taftp, weights, out, nantennas_total=None, start_antenna=0,
                       nsamples_per_heap=256, stream=None):
        """Coherent beams straight from a TAFTP block (the plan's antennas
        start at start_antenna of the block's nantennas_total)."""
        if nantennas_total is None:
            nantennas_total = self.desc.nantennas
        call("fbf_cb_beamform_taftp", self.handle, taftp.ptr, taftp.nbytes, int(nantennas_total),
             int(start_antenna), int(nsamples_per_heap), weights.ptr, weights.nbytes, out.ptr, out.nbytes,
             _s(stream))

    def ib_beamform_taftp(self, ib_plan, taftp, weights, out, tf, nsamples_per_heap=256, stream=None):
        """Coherent beams (out) and the incoherent beam of ib_plan (tf) of a
        TAFTP block of this plan's antennas, in one pass where the plans
        allow it (ib_fused)."""
        call("fbf_cb_ib_beamform_taftp", self.handle, ib_plan.handle, taftp.ptr, taftp.nbytes,
             int(nsamples_per_heap), weights.ptr, weights.nbytes, out.ptr, out.nbytes, tf.ptr, tf.nbytes,
             _s(stream))

    def ib_beamform_taftp_next(self, ib_plan, taftp, weights, out, tf, weights_plan, delays, next_epoch,
                               next_weights, nsamples_per_heap=256, stream=None):
        """ib_beamform_taftp, then the next block's weights (at next_epoch,
        into next_weights, which may be `weights`)."""
        call("fbf_cb_ib_beamform_taftp_next", self.handle, ib_plan.handle, taftp.ptr, taftp.nbytes,
             int(nsamples_per_heap), weights.ptr, weights.nbytes, out.ptr, out.nbytes, tf.ptr, tf.nbytes,
             weights_plan.handle, delays.ptr, delays.nbytes, float(next_epoch), next_weights.ptr,
             next_weights.nbytes, _s(stream))

    def ib_fused(self, ib_plan, nsamples, nsamples_per_heap=256):
        """Whether ib_beamform_taftp forms the incoherent beam inside the
        coherent kernel."""
        f = C.c_int()
        call("fbf_cb_ib_fused", self.handle, ib_plan.handle, int(nsamples_per_heap), int(nsamples), C.byref(f))
        return bool(f.value)

    def taftp_direct(self, nantennas_total, nsamples, nsamples_per_heap=256):
        """Whether beamform_taftp reads such a block without an FTPA copy."""
        d = C.c_int()
        call("fbf_cb_taftp_direct", self.handle, int(nantennas_total), int(nsamples_per_heap), int(nsamples),
             C.byref(d))
        return bool(d.value)

    def beam_sums(self, ftpa, weights, out, stream=None):
        call("fbf_cb_beam_sums", self.handle, ftpa.ptr, ftpa.nbytes, weights.ptr, weights.nbytes,
             out.ptr, out.nbytes, _s(stream))

    def beam_power(self, ftpa, weights, out, stream=None):
        call("fbf_cb_beam_power", self.handle, ftpa.ptr, ftpa.nbytes, weights.ptr, weights.nbytes,
             out.ptr, out.nbytes, _s(stream))

    def close(self):
        if self.handle:
            call("fbf_cb_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IbPlan:
    """Incoherent beamformer plan (fbf_ib_plan)."""

    def __init__(self, nantennas, nchans, tscrunch, fscrunch, offsets, scales,
                 nsamples_per_heap=256, quant_mode=QUANT_ROUND):
        self.desc = IbDesc(nantennas, nchans, nsamples_per_heap, tscrunch, fscrunch, quant_mode)
        off = np.ascontiguousarray(offsets, dtype=np.float32)
        sc = np.ascontiguousarray(scales, dtype=np.float32)
        h = C.c_void_p()
        call("fbf_ib_plan_create", C.byref(self.desc), _np_ptr(off), _np_ptr(sc), C.byref(h))
        self.handle = h

    def output_bytes(self, taftp_bytes):
        n = C.c_size_t()
        call("fbf_ib_output_bytes", self.handle, taftp_bytes, C.byref(n))
        return n.value

    def beamform(self, taftp, out, stream=None):
        call("fbf_ib_beamform", self.handle, taftp.ptr, taftp.nbytes, out.ptr, out.nbytes, _s(stream))

    def beamform_weights(self, taftp, out, weights_plan, delays, epoch, weights, stream=None):
        """beamform and the weights generator in one launch."""
        call("fbf_ib_beamform_weights", self.handle, taftp.ptr, taftp.nbytes, out.ptr, out.nbytes,
             weights_plan.handle, delays.ptr, delays.nbytes, float(epoch), weights.ptr, weights.nbytes,
             _s(stream))

    def power(self, taftp, out, stream=None):
        call("fbf_ib_power", self.handle, taftp.ptr, taftp.nbytes, out.ptr, out.nbytes, _s(stream))

    def close(self):
        if self.handle:
            call("fbf_ib_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WeightsPlan:
    """Delay -> weights generator plan (fbf_weights_plan)."""

    def __init__(self, nantennas, nbeams, nchans, chan_freqs):
        fr = np.ascontiguousarray(chan_freqs, dtype=np.float64)
        assert fr.size == nchans
        h = C.c_void_p()
        call("fbf_weights_plan_create", nantennas, nbeams, nchans, _np_ptr(fr), C.byref(h))
        self.handle = h

    def generate(self, delays, epoch, weights, phase=None, stream=None):
        call("fbf_weights_generate", self.handle, delays.ptr, delays.nbytes, float(epoch),
             weights.ptr, weights.nbytes, phase.ptr if phase is not None else None, _s(stream))

    def generate_steps(self, delays, tstart, tstep, ntsteps, weights, stream=None):
        """Weights of epochs tstart + k * tstep, k < ntsteps, as [t][F][B][A]."""
        call("fbf_weights_generate_steps", self.handle, delays.ptr, delays.nbytes, float(tstart), float(tstep),
             int(ntsteps), weights.ptr, weights.nbytes, _s(stream))

    def close(self):
        if self.handle:
            call("fbf_weights_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched subarrays on one block (fbf_batch)."""

    def __init__(self, total_nantennas, nchans, nheap_groups, subarrays, levels, nsamples_per_heap=256):
        """subarrays: list of dicts (start_antenna, nantennas, nbeams, tscrunch, fscrunch,
        [nsamples_per_heap_out], [quant_mode]); levels: list of (offsets, scales)."""
        n = len(subarrays)
        arr = (SubarrayDesc * n)()
        for i, sd in enumerate(subarrays):
            arr[i] = SubarrayDesc(sd["start_antenna"], sd["nantennas"], sd["nbeams"], sd["tscrunch"],
                                  sd["fscrunch"], sd.get("nsamples_per_heap_out", 0), sd.get("quant_mode", 0))
        self._levels = [(np.ascontiguousarray(o, np.float32), np.ascontiguousarray(s, np.float32))
                        for o, s in levels]
        offs = (C.c_void_p * n)(*[_np_ptr(o).value for o, _ in self._levels])
        scs = (C.c_void_p * n)(*[_np_ptr(s).value for _, s in self._levels])
        h = C.c_void_p()
        call("fbf_batch_create", total_nantennas, nchans, nsamples_per_heap, nheap_groups, n, arr, offs, scs,
             C.byref(h))
        self.handle = h
        self.n = n
        self.ib = False

    def enable_ib(self, tscrunch, fscrunch, offsets, scales, quant_mode=QUANT_ROUND):
        o = np.ascontiguousarray(offsets, np.float32)
        s = np.ascontiguousarray(scales, np.float32)
        call("fbf_batch_enable_ib", self.handle, tscrunch, fscrunch, quant_mode, _np_ptr(o), _np_ptr(s))
        self.ib = True

    def output_bytes(self, i):
        b = C.c_size_t()
        call("fbf_batch_output_bytes", self.handle, i, C.byref(b))
        return b.value

    def weights_bytes(self, i):
        b = C.c_size_t()
        call("fbf_batch_weights_bytes", self.handle, i, C.byref(b))
        return b.value

    def set_variant(self, i, variant, chunk=0):
        call("fbf_batch_set_variant", self.handle, i, variant, chunk)

    def autotune(self, reps=3):
        call("fbf_batch_autotune", self.handle, reps)

    def set_serial(self, serial):
        """1: every kernel of a block in turn on the caller's stream; 0: subarrays side by side."""
        call("fbf_batch_set_serial", self.handle, int(bool(serial)))

    @property
    def serial(self):
        v = C.c_int()
        call("fbf_batch_get_serial", self.handle, C.byref(v))
        return bool(v.value)

    def run(self, taftp, weights, outputs, use_graph=True, stream=None):
        w = (C.c_void_p * self.n)(*[x.ptr.value for x in weights])
        o = (C.c_void_p * len(outputs))(*[x.ptr.value for x in outputs])
        call("fbf_batch_run", self.handle, taftp.ptr, taftp.nbytes, w, o, int(bool(use_graph)), _s(stream))

    def close(self):
        if self.handle:
            call("fbf_batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Pinned host allocation exposed as a numpy array."""

    def __init__(self, nbytes, dtype=np.int8):
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        call("fbf_host_malloc", C.byref(h), self.nbytes)
        self.ptr = h
        dtype = np.dtype(dtype)
        buf = (C.c_char * self.nbytes).from_address(h.value)
        self.array = np.frombuffer(buf, dtype=dtype, count=self.nbytes // dtype.itemsize)

    def free(self):
        if self.ptr:
            self.array = None
            call("fbf_host_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Serving engine (fbf_server): feeds, continuous batching, latency stats."""

    def __init__(self, nslots, hbm_budget):
        h = C.c_void_p()
        call("fbf_server_create", int(nslots), int(hbm_budget), C.byref(h))
        self.handle = h

    def add_feed(self, nantennas, nbeams, nchans, nheap_groups, tscrunch, fscrunch, cb_levels,
                 ib=None, ib_levels=None, nsamples_per_heap_out=0, quant_mode=QUANT_ROUND):
        d = FeedDesc(nantennas, nbeams, nchans, nheap_groups, 256, tscrunch, fscrunch, nsamples_per_heap_out,
                     quant_mode, ib[0] if ib else 0, ib[1] if ib else 0)
        co = np.ascontiguousarray(cb_levels[0], np.float32)
        cs = np.ascontiguousarray(cb_levels[1], np.float32)
        io = np.ascontiguousarray(ib_levels[0], np.float32) if ib else None
        is_ = np.ascontiguousarray(ib_levels[1], np.float32) if ib else None
        f = C.c_int()
        call("fbf_server_add_feed", self.handle, C.byref(d), _np_ptr(co), _np_ptr(cs),
             _np_ptr(io) if io is not None else None, _np_ptr(is_) if is_ is not None else None, C.byref(f))
        return f.value

    def feed_bytes(self, feed):
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        call("fbf_server_feed_bytes", self.handle, feed, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def set_weights(self, feed, weights):
        w = np.ascontiguousarray(weights)
        call("fbf_server_set_weights", self.handle, feed, _np_ptr(w), w.nbytes)

    def set_delays(self, feed, delays, chan_freqs, epoch):
        d = np.ascontiguousarray(delays, np.float32)
        fr = np.ascontiguousarray(chan_freqs, np.float64)
        call("fbf_server_set_delays", self.handle, feed, _np_ptr(d), d.nbytes, _np_ptr(fr), float(epoch))

    def start(self):
        call("fbf_server_start", self.handle)

    def submit(self, feed, inp, cb_out, ib_out=None, tag=0, on_device=False):
        """inp: HostBuffer or DeviceBuffer (on_device=True); cb_out/ib_out: HostBuffer."""
        call("fbf_server_submit", self.handle, feed, inp.ptr, inp.nbytes, int(bool(on_device)), cb_out.ptr,
             ib_out.ptr if ib_out is not None else None, C.c_uint64(tag))

    def poll(self, timeout_ms=1000):
        t, st, lat = C.c_uint64(), C.c_int(), C.c_double()
        r = lib().fbf_server_poll(self.handle, C.byref(t), C.byref(st), C.byref(lat), int(timeout_ms))
        if r == FBF_ERR_STATE:
            return None
        check(r)
        return t.value, st.value, lat.value

    def stats(self):
        st = ServerStats()
        call("fbf_server_stats_get", self.handle, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        call("fbf_server_reset_stats", self.handle)

    def drain(self):
        call("fbf_server_drain", self.handle)

    def close(self):
        if self.handle:
            call("fbf_server_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
