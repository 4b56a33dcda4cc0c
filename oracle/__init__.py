"""CPU oracle for the FBFUSE beamformer path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
(``psrdada_cpp_amd``) never imports it and has no CPU fallback.

Thin numpy wrappers over ``libfbf_oracle.so`` (built from ``fbf_oracle.c`` by
``oracle/Makefile``).  Each function cites the reference file:line whose
behaviour it restates (paths relative to the psrdada_cpp source tree).

Pinning (what checks each function beyond restating it):
  split_transpose  -- hand-derived index KAT, permutation invariant, brute-force
                      loop in tests/test_oracle.py
  cb_sums          -- closed forms (unit weights, single antenna, quarter-turn
                      phases), linearity, numpy complex matmul
  cb_power         -- numpy complex matmul + |.|^2 + scrunch, closed form
  quantise         -- hand-computed edge cases (+-0.5, saturation, truncation)
  cb_reference_f32 -- agrees with the exact path within the reference's +-1
  ib_power         -- closed form for constant input, numpy brute force
  weights_f64      -- zero delay -> (127, 0), tau -> -tau conjugates,
                      integer cycles -> (127, 0), numpy FP64
  power_levels     -- closed-form formulas of PipelineConfig.cpp:109-127
  channel_freqs    -- closed form of PipelineConfig.cpp:180-195
  tuse_transpose   -- hand-derived index KAT, numpy reshape/transpose,
                      permutation invariant (per-beam byte histogram)
  feng_to_dada     -- hand-derived index KAT, numpy transpose, involution
                      with the inverse reorder
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfbf_oracle.so")

QUANT_ROUND = 0
QUANT_TRUNC = 1

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library {_LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i8(a):
    a = np.ascontiguousarray(a)
    assert a.dtype == np.int8, a.dtype
    return a


def split_transpose(taftp, nheap_groups, total_nantennas, used_nantennas,
                    start_antenna, nchans, nsamples_per_heap=256, npol=2):
    """SplitTransposeTester.cu:33-75. Bytes in, bytes out (char2 = 2 bytes)."""
    taftp = _i8(taftp)
    out = np.empty(2 * nchans * nheap_groups * nsamples_per_heap * npol * used_nantennas,
                   dtype=np.int8)
    _load().oracle_split_transpose(_p(taftp), _p(out), ctypes.c_int(nheap_groups),
                                   ctypes.c_int(total_nantennas), ctypes.c_int(used_nantennas),
                                   ctypes.c_int(start_antenna), ctypes.c_int(nchans),
                                   ctypes.c_int(nsamples_per_heap), ctypes.c_int(npol))
    return out


def cb_sums(ftpa, weights, nchans, nsamples, nantennas, nbeams, npol=2, nthreads=8):
    """CoherentBeamformerTester.cu:75-96: int32 [F][T][P][B][2]."""
    ftpa, weights = _i8(ftpa), _i8(weights)
    out = np.empty((nchans, nsamples, npol, nbeams, 2), dtype=np.int32)
    _load().oracle_cb_sums(_p(ftpa), _p(weights), _p(out), ctypes.c_int(nchans),
                           ctypes.c_long(nsamples), ctypes.c_int(npol),
                           ctypes.c_int(nantennas), ctypes.c_int(nbeams),
                           ctypes.c_int(nthreads))
    return out


def cb_power(ftpa, weights, nchans, nsamples, nantennas, nbeams, tscrunch, fscrunch,
             npol=2, nthreads=8):
    """Exact power [F/fs][T/ts][B] (CoherentBeamformerTester.cu:55-100)."""
    ftpa, weights = _i8(ftpa), _i8(weights)
    out = np.empty((nchans // fscrunch, nsamples // tscrunch, nbeams), dtype=np.float64)
    _load().oracle_cb_power(_p(ftpa), _p(weights), _p(out), ctypes.c_int(nchans),
                            ctypes.c_long(nsamples), ctypes.c_int(npol),
                            ctypes.c_int(nantennas), ctypes.c_int(nbeams),
                            ctypes.c_int(tscrunch), ctypes.c_int(fscrunch),
                            ctypes.c_int(nthreads))
    return out


def quantise(power, offset, scale, mode=QUANT_ROUND):
    lib = _load()
    lib.oracle_quantise.restype = ctypes.c_int8
    return int(lib.oracle_quantise(ctypes.c_double(power), ctypes.c_float(offset),
                                   ctypes.c_float(scale), ctypes.c_int(mode)))


def cb_tbtf(power, offsets, scales, nsamples_per_heap_out, mode=QUANT_ROUND):
    """Power [Fo][To][B] -> int8 TBTF (CoherentBeamformer.cu:162-169)."""
    power = np.ascontiguousarray(power, dtype=np.float64)
    fo, to, nb = power.shape
    offsets = np.ascontiguousarray(offsets, dtype=np.float32)
    scales = np.ascontiguousarray(scales, dtype=np.float32)
    out = np.empty(fo * to * nb, dtype=np.int8)
    _load().oracle_cb_tbtf(_p(power), _p(out), _p(offsets), _p(scales), ctypes.c_int(nb),
                           ctypes.c_int(fo), ctypes.c_long(to),
                           ctypes.c_int(nsamples_per_heap_out), ctypes.c_int(mode))
    return out


def cb_reference_f32(ftpa, weights, nchans, tscrunch, fscrunch, nsamples, nbeams,
                     nantennas, scale, offset, nsamples_per_heap_out, npol=2):
    """CoherentBeamformerTester.cu:37-118 in FP32, loop for loop."""
    ftpa, weights = _i8(ftpa), _i8(weights)
    out = np.empty(nchans // fscrunch * nsamples // tscrunch * nbeams, dtype=np.int8)
    _load().oracle_cb_reference_f32(_p(ftpa), _p(weights), _p(out), ctypes.c_int(nchans),
                                    ctypes.c_int(tscrunch), ctypes.c_int(fscrunch),
                                    ctypes.c_long(nsamples), ctypes.c_int(nbeams),
                                    ctypes.c_int(nantennas), ctypes.c_int(npol),
                                    ctypes.c_float(scale), ctypes.c_float(offset),
                                    ctypes.c_int(nsamples_per_heap_out))
    return out


def ib_power(taftp, nchans, tscrunch, fscrunch, ntimestamps, nantennas,
             nsamples_per_heap=256, npol=2, nthreads=8):
    """IncoherentBeamformerTester.cu:36-97, exact: double [hg*S/ts][F/fs]."""
    taftp = _i8(taftp)
    out = np.empty((ntimestamps * nsamples_per_heap // tscrunch, nchans // fscrunch),
                   dtype=np.float64)
    _load().oracle_ib_power(_p(taftp), _p(out), ctypes.c_int(nchans), ctypes.c_int(tscrunch),
                            ctypes.c_int(fscrunch), ctypes.c_int(ntimestamps),
                            ctypes.c_int(nantennas), ctypes.c_int(npol),
                            ctypes.c_int(nsamples_per_heap), ctypes.c_int(nthreads))
    return out


def ib_tf(power, offsets, scales, mode=QUANT_ROUND):
    power = np.ascontiguousarray(power, dtype=np.float64)
    nt, nc = power.shape
    offsets = np.ascontiguousarray(offsets, dtype=np.float32)
    scales = np.ascontiguousarray(scales, dtype=np.float32)
    out = np.empty(nt * nc, dtype=np.int8)
    _load().oracle_ib_tf(_p(power), _p(out), _p(offsets), _p(scales), ctypes.c_long(nt),
                         ctypes.c_int(nc), ctypes.c_int(mode))
    return out


def ib_reference_f32(taftp, nchans, tscrunch, fscrunch, ntimestamps, nantennas,
                     scale, offset, nsamples_per_heap=256, npol=2):
    taftp = _i8(taftp)
    out = np.empty(ntimestamps * nsamples_per_heap // tscrunch * (nchans // fscrunch),
                   dtype=np.int8)
    _load().oracle_ib_reference_f32(_p(taftp), _p(out), ctypes.c_int(nchans),
                                    ctypes.c_int(tscrunch), ctypes.c_int(fscrunch),
                                    ctypes.c_int(ntimestamps), ctypes.c_int(nantennas),
                                    ctypes.c_int(npol), ctypes.c_int(nsamples_per_heap),
                                    ctypes.c_float(scale), ctypes.c_float(offset))
    return out


def weights_f64(delays, freqs, epoch, nantennas, nbeams, nchans, want_phase=False):
    """FP64 weights [F][B][A] char2 (WeightsManager.cu:58-73)."""
    delays = np.ascontiguousarray(delays, dtype=np.float32)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.empty(2 * nchans * nbeams * nantennas, dtype=np.int8)
    phase = np.empty(nchans * nbeams * nantennas, dtype=np.float64) if want_phase else None
    _load().oracle_weights_f64(_p(delays), _p(freqs), ctypes.c_double(epoch),
                               ctypes.c_int(nantennas), ctypes.c_int(nbeams),
                               ctypes.c_int(nchans), _p(out),
                               _p(phase) if phase is not None else ctypes.c_void_p(0))
    return (out, phase) if want_phase else out


def weights_reference_f32(delays, freqs, epoch, nantennas, nbeams, nchans):
    """WeightsManagerTester.cu:38-73 in FP32."""
    delays = np.ascontiguousarray(delays, dtype=np.float32)
    freqs = np.ascontiguousarray(freqs, dtype=np.float32)
    out = np.empty(2 * nchans * nbeams * nantennas, dtype=np.int8)
    _load().oracle_weights_reference_f32(_p(delays), _p(freqs), ctypes.c_float(epoch),
                                         ctypes.c_int(nantennas), ctypes.c_int(nbeams),
                                         ctypes.c_int(nchans), _p(out))
    return out


def power_levels(input_level, output_level, cb_nantennas, cb_tscrunch, cb_fscrunch,
                 ib_nantennas, ib_tscrunch, ib_fscrunch, npol=2):
    """PipelineConfig.cpp:109-127 -> (cb_offset, cb_scale, ib_offset, ib_scale)."""
    vals = [ctypes.c_float() for _ in range(4)]
    _load().oracle_power_levels(ctypes.c_float(input_level), ctypes.c_float(output_level),
                                ctypes.c_int(cb_nantennas), ctypes.c_int(cb_tscrunch),
                                ctypes.c_int(cb_fscrunch), ctypes.c_int(ib_nantennas),
                                ctypes.c_int(ib_tscrunch), ctypes.c_int(ib_fscrunch),
                                ctypes.c_int(npol), *[ctypes.byref(v) for v in vals])
    return tuple(v.value for v in vals)


def channel_frequencies(cfreq, bw, nchans):
    """PipelineConfig.cpp:180-195 (float)."""
    out = np.empty(nchans, dtype=np.float32)
    _load().oracle_channel_frequencies(ctypes.c_float(cfreq), ctypes.c_float(bw),
                                       ctypes.c_int(nchans), _p(out))
    return out


def tuse_transpose(block, nchans, nsamples, nfreq, nbeams, ngroups, beam0=0, nbeams_out=None):
    """TUSE corner turn, per beam [ngroups][nsamples][nfreq][nchans]
    (meerkat/tuse/src/transpose_to_dada.cpp:19-48)."""
    block = _i8(block)
    if nbeams_out is None:
        nbeams_out = nbeams - beam0
    assert block.size >= ngroups * nbeams * nfreq * nsamples * nchans
    out = np.empty(nbeams_out * ngroups * nsamples * nfreq * nchans, np.int8)
    _load().oracle_tuse_transpose(_p(block), _p(out), ctypes.c_int(nchans), ctypes.c_int(nsamples),
                                  ctypes.c_int(nfreq), ctypes.c_int(nbeams), ctypes.c_int(ngroups),
                                  ctypes.c_int(beam0), ctypes.c_int(nbeams_out))
    return out


def feng_to_dada(words, nchans):
    """F-engine heap reorder, per timestamp [nchans][256] -> [256][nchans]
    32-bit words (meerkat/tools/src/feng_to_dada.cu:54-80)."""
    words = np.ascontiguousarray(words, dtype=np.int32)
    assert words.size % (nchans * 256) == 0
    out = np.empty_like(words)
    _load().oracle_feng_to_dada(_p(words), _p(out), ctypes.c_int(nchans),
                                ctypes.c_long(words.size // (nchans * 256)))
    return out
