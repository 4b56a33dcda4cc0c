"""CPU-side checks of the native library: it loads, exports every symbol the
public headers declare, and its host-only helpers agree with the oracle.
No compute entry point is called here (no GPU in this environment)."""
import ctypes

import numpy as np
import pytest


def test_library_exports_every_header_symbol(capi):
    names = capi.header_symbols()
    assert len(names) > 40
    L = capi.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_power_levels_match_oracle(capi, oracle):
    for args in [(32.0, 24.0, 44, 16, 4, 64, 16, 2), (16.0, 24.0, 64, 16, 1, 64, 16, 1),
                 (16.0, 24.0, 8, 64, 1, 64, 16, 1), (7.5, 3.0, 40, 4, 2, 40, 4, 2)]:
        assert capi.power_levels(*args) == oracle.power_levels(*args)


def test_channel_frequencies_match_oracle(capi, oracle):
    for cf, bw, n in [(1.4e9, 56e6, 64), (1284e6, 856e6, 16), (1284e6, 13.375e6, 64)]:
        np.testing.assert_array_equal(capi.channel_frequencies(cf, bw, n), oracle.channel_frequencies(cf, bw, n))
    f64 = capi.channel_frequencies_f64(1284e6, 856e6, 16)
    np.testing.assert_allclose(f64, 856e6 + 53.5e6 * (np.arange(16) + 0.5), rtol=0, atol=1e-3)


def test_synth_gaussian_statistics_and_determinism(capi):
    a = capi.synth_gaussian_int8(123, 16.0, 1 << 20)
    b = capi.synth_gaussian_int8(123, 16.0, 1 << 20)
    np.testing.assert_array_equal(a, b)
    c = capi.synth_gaussian_int8(124, 16.0, 1 << 20)
    assert np.mean(a != c) > 0.9
    x = a.astype(np.float64)
    assert abs(x.mean()) < 0.1 and abs(x.std() - 16.0) < 0.1
    assert a.min() >= -127 and a.max() <= 127
    # prefix property (counter-based stream): the first n bytes do not depend on the total length
    np.testing.assert_array_equal(capi.synth_gaussian_int8(123, 16.0, 1001), a[:1001])


def test_synth_wraps_without_clip(capi):
    a = capi.synth_gaussian_int8(5, 100.0, 1 << 16, clip=False)
    assert a.min() == -128   # values beyond the int8 range wrap like static_cast<int8_t>


def test_synth_uniform(capi):
    u = capi.synth_uniform_f32(9, -1e-6, 1e-6, 100000)
    assert u.min() >= -1e-6 and u.max() < 1e-6 and abs(u.mean()) < 1e-8


def test_argument_errors_are_status_codes(capi):
    """Validation happens on the host before any HIP call."""
    L = capi.lib()
    st = L.fbf_split_transpose(None, 0, None, 0, 64, 44, 0, 64, 256, None)
    assert st == capi.FBF_ERR_ARG
    assert b"null" in L.fbf_last_error()
    # wrong output size -> FBF_ERR_SIZE (no launch happens)
    buf = ctypes.create_string_buffer(64)
    st = L.fbf_split_transpose(ctypes.cast(buf, ctypes.c_void_p), 4 * 64 * 64 * 256,
                               ctypes.cast(buf, ctypes.c_void_p), 12, 64, 44, 0, 64, 256, None)
    assert st == capi.FBF_ERR_SIZE
    desc = capi.CbDesc(44, 1020, 64, 16, 4, 0, 0)   # 1020 beams: not a multiple of 8
    h = ctypes.c_void_p()
    z = np.zeros(16, np.float32)
    st = L.fbf_cb_plan_create(ctypes.byref(desc), z.ctypes.data_as(ctypes.c_void_p),
                              z.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))
    assert st == capi.FBF_ERR_UNSUPPORTED
    with pytest.raises(capi.FbfError):
        capi.check(st)
