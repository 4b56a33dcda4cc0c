"""fbf_cb_beamform_noalloc, the coherent-beam entry point that allocates
nothing (its staging buffer is a fixed device array of the library), against
fbf_cb_beamform through the C ABI: two calls in a row, the second with a
smaller input, two plans of different geometry sharing the array, and the
size limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import gaussian_int8, unit_weights

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def _noalloc(capi):
    fn = capi.lib().fbf_cb_beamform_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p]
    return fn


def _levels(oracle, A, ts, F):
    cbo, cbs, _, _ = oracle.power_levels(16.0, 24.0, A, ts, 1, 64, 16, 1)
    return np.full(F, cbo, np.float32), np.full(F, cbs, np.float32)


def test_noalloc_equals_beamform_twice_second_smaller(gpu, oracle):
    capi = gpu
    fn = _noalloc(capi)
    F, A, B, ts = 64, 64, 512, 16
    off, sc = _levels(oracle, A, ts, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc)
    # a second plan of another geometry shares the same staging array
    F2, A2, B2 = 16, 40, 96
    off2, sc2 = _levels(oracle, A2, ts, F2)
    plan2 = capi.CbPlan(A2, B2, F2, ts, 1, off2, sc2, nsamples_per_heap_out=64)
    s = capi.Stream()
    for call, T in enumerate((8192, 2048)):
        print(f"noalloc call {call}: {F} ch x {T} samples, {A} ant, {B} beams", flush=True)
        ftpa = gaussian_int8(70 + call, F * T * 2 * A * 2, 16.0)
        w = unit_weights(80 + call, F * B * A)
        d_v, d_w = capi.DeviceBuffer.from_array(ftpa), capi.DeviceBuffer.from_array(w)
        n = plan.output_bytes(ftpa.nbytes)
        d_ref, d_got = capi.DeviceBuffer(n), capi.DeviceBuffer(n)
        d_got.fill(0x5A, s)
        plan.beamform(d_v, d_w, d_ref, stream=s)
        capi.check(fn(plan.handle, d_v.ptr, d_v.nbytes, d_w.ptr, d_w.nbytes, d_got.ptr, d_got.nbytes, s.handle))
        s.sync()
        np.testing.assert_array_equal(d_got.download(), d_ref.download(), err_msg=f"call {call}")
        # the other plan between the calls
        T2 = 1024
        v2 = gaussian_int8(90 + call, F2 * T2 * 2 * A2 * 2, 16.0)
        w2 = unit_weights(91 + call, F2 * B2 * A2)
        d_v2, d_w2 = capi.DeviceBuffer.from_array(v2), capi.DeviceBuffer.from_array(w2)
        n2 = plan2.output_bytes(v2.nbytes)
        d_o2 = capi.DeviceBuffer(n2)
        capi.check(fn(plan2.handle, d_v2.ptr, d_v2.nbytes, d_w2.ptr, d_w2.nbytes, d_o2.ptr, d_o2.nbytes, s.handle))
        s.sync()
        p2 = oracle.cb_power(v2, w2, F2, T2, A2, B2, ts, 1)
        np.testing.assert_array_equal(d_o2.download(), oracle.cb_tbtf(p2, off2, sc2, 64))
    # the smaller call also against the oracle
    p = oracle.cb_power(ftpa, w, F, 2048, A, B, ts, 1, nthreads=16)
    np.testing.assert_array_equal(d_got.download(), oracle.cb_tbtf(p, off, sc, 8192 // F))


def test_noalloc_rejects_outputs_above_its_array(gpu):
    """An output larger than the fixed staging array is refused with a size
    error before anything is launched (only the sizes are read)."""
    capi = gpu
    fn = _noalloc(capi)
    F, A, B, ts = 64, 64, 512, 16
    plan = capi.CbPlan(A, B, F, ts, 1, np.ones(F, np.float32), np.ones(F, np.float32))
    d = capi.DeviceBuffer(4096)
    To = SCRATCH_BYTES // (B * F) + 128       # whole output heaps of 128 samples
    ftpa_bytes = To * ts * F * 2 * A * 2
    out_bytes = To * B * F
    st = fn(plan.handle, d.ptr, ftpa_bytes, d.ptr, F * B * A * 2, d.ptr, out_bytes, None)
    assert st == capi.FBF_ERR_SIZE, (st, capi.lib().fbf_last_error())
    assert b"staging array" in capi.lib().fbf_last_error()
