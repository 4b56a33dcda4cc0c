"""The coherent beamformer's quantiser (cb_tr_k) replaces the division of the
exact quantiser by a per-channel table: thr[j] is the least integer power
that quantises to at least j - 128.  An estimate e = trunc(clamp(y)) of the
quotient, within one step of the result in either rounding mode, is
corrected by the two thresholds around it.  This checks that arithmetic on
the CPU (numpy float64 and float32 are IEEE, as on the device) against the
oracle's quantiser, with thresholds found by binary search on the oracle
itself; the GPU tests check the kernel.
"""
import numpy as np
import pytest

INT64_MIN, INT64_MAX = np.iinfo(np.int64).min, np.iinfo(np.int64).max


def _thresholds(oracle, off, sc, mode):
    t = np.empty(257, np.int64)
    t[0], t[256] = INT64_MIN, INT64_MAX
    for j in range(1, 256):
        k = j - 128
        lo, hi = 0, 1 << 62
        if oracle.quantise(float(hi), off, sc, mode) < k:
            t[j] = INT64_MAX
            continue
        while lo < hi:
            mid = lo + (hi - lo) // 2
            if oracle.quantise(float(mid), off, sc, mode) >= k:
                hi = mid
            else:
                lo = mid + 1
        t[j] = lo
    return t


def _settle(P, e, thr):
    """Threshold correction of the estimate e in [-128, 127]."""
    j = e.astype(np.int64) + 128
    j += (P >= thr[j + 1]).astype(np.int64) - (P < thr[j]).astype(np.int64)
    return (j - 128).astype(np.int8)


def _quantise_thr(P, off, sc, thr, mode):
    """FP64 estimate RN(RN(P - offset) * RN(1 / scale)), truncated whatever
    the mode (the correction absorbs the rounding)."""
    y = (P.astype(np.float64) - np.float64(np.float32(off))) * (1.0 / np.float64(np.float32(sc)))
    return _settle(P, np.trunc(np.clip(y, -128, 127)), thr)


# (offset, scale): config-2 style levels, levels whose quotients are exact
# integers and half-integers, large and tiny scales, a negative offset
LEVELS = [(1.41e9, 2.6e7), (3.5, 1.0), (12.0, 8.0), (6.0e4, 2048.0), (7.0e12, 3.1e10),
          (-2.5e5, 0.37), (0.0, 1e-3), (5.0e15, 1.0e13), (4194304.5, 1.0), (3.0e9, 7.0)]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("off,sc", LEVELS)
def test_threshold_quantiser_matches_exact_division(oracle, off, sc, mode):
    thr = _thresholds(oracle, off, sc, mode)
    assert np.all(thr[1:] >= thr[:-1])
    reach = thr[1:256][thr[1:256] < INT64_MAX]
    # powers at and next to every threshold, and a spread over 2^0 .. 2^56
    near = (reach[:, None] + np.arange(-3, 4)[None, :]).reshape(-1)
    rng = np.random.default_rng(3)
    spread = (rng.integers(0, 1 << 56, 20000) >> rng.integers(0, 56, 20000)).astype(np.int64)
    P = np.concatenate([near[near >= 0], spread])
    got = _quantise_thr(P, off, sc, thr, mode)
    want = oracle.cb_tbtf(P.astype(np.float64).reshape(1, 1, -1), np.float32([off]), np.float32([sc]), 1,
                          mode).reshape(-1)
    np.testing.assert_array_equal(got, want)
    # the levels really cross steps: most thresholds are reachable
    assert reach.size > 100


def _quantise_thr_f32(P, off, sc, thr, mode):
    """The FP32 estimate of the coherent beamformer (used when
    |offset / scale| < 2^20): fma(pf, RN(1 / scale), RN(-offset * inv)),
    truncated, corrected by the same thresholds.  The fma is modelled by the
    exact product in float64 plus the constant, rounded to float64 and then
    to float32 (one rounding more than the device's fma)."""
    f32 = np.float32
    pf = (P >> 32).astype(f32) * f32(4294967296.0) + (P & 0xFFFFFFFF).astype(f32)
    inv = f32(1.0) / f32(sc)
    c = -f32(off) * inv
    y = (pf.astype(np.float64) * np.float64(inv) + np.float64(c)).astype(f32)
    return _settle(P, np.trunc(np.clip(y, -128, 127)), thr)


# |offset / scale| below 2^20, up to the bound
LEVELS_F32 = [(1.41e9, 2.6e7), (3.5, 1.0), (12.0, 8.0), (-2.5e5, 0.37), (1.0e6 * 3.0, 3.0),
              (-9.9e5 * 0.125, 0.125), (6.0e4, 2048.0)]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("off,sc", LEVELS_F32)
def test_threshold_quantiser_f32_estimate_is_exact(oracle, off, sc, mode):
    assert abs(np.float32(off) / np.float32(sc)) < 2 ** 20
    thr = _thresholds(oracle, off, sc, mode)
    reach = thr[1:256][thr[1:256] < INT64_MAX]
    near = (reach[:, None] + np.arange(-3, 4)[None, :]).reshape(-1)
    rng = np.random.default_rng(9)
    spread = (rng.integers(0, 1 << 62, 20000) >> rng.integers(0, 62, 20000)).astype(np.int64)
    P = np.concatenate([near[near >= 0], spread])
    want = oracle.cb_tbtf(P.astype(np.float64).reshape(1, 1, -1), np.float32([off]), np.float32([sc]), 1,
                          mode).reshape(-1)
    np.testing.assert_array_equal(_quantise_thr_f32(P, off, sc, thr, mode), want)
    # the estimate's error bound, 2^-24 (651 + 6 |offset / scale|) for
    # |Y| <= 130, stays below 0.5 (so the truncated estimate is within a step)
    f32 = np.float32
    o = abs(float(f32(off)) / float(f32(sc)))
    pf = (P >> 32).astype(f32) * f32(4294967296.0) + (P & 0xFFFFFFFF).astype(f32)
    inv = f32(1.0) / f32(sc)
    y = (pf.astype(np.float64) * np.float64(inv) + np.float64(-f32(off) * inv)).astype(f32).astype(np.float64)
    Y = (P.astype(np.float64) - float(f32(off))) / float(f32(sc))
    near = np.abs(Y) <= 130     # empty for levels where every P >= 0 saturates
    bound = 2.0 ** -24 * (651 + 6 * o)
    assert bound < 0.5
    assert np.all(np.abs(y - Y)[near] <= bound)


def test_threshold_estimate_alone_is_not_exact(oracle):
    """Without the threshold correction the estimate disagrees with the
    exact quantiser at some boundary (the correction is not idle)."""
    off, sc, mode = 3.5, 1.0, 0
    thr = _thresholds(oracle, off, sc, mode)
    P = np.arange(0, 300, dtype=np.int64)
    y = (P.astype(np.float64) - 3.5) / 1.0
    est = np.clip(np.rint(y), -128, 127).astype(np.int8)   # half-even, not half-away
    want = oracle.cb_tbtf(P.astype(np.float64).reshape(1, 1, -1), np.float32([off]), np.float32([sc]), 1,
                          mode).reshape(-1)
    assert np.any(est != want)
    np.testing.assert_array_equal(_quantise_thr(P, off, sc, thr, mode), want)


def test_fp64_estimate_needs_the_difference_first(oracle):
    """The FP64 estimate rounds (P - offset) before scaling, as the exact
    quantiser does, so its error is relative to the small quotient.  The
    single-fma form P * inv - offset * inv (right for the FP32 path, whose
    |offset / scale| is bounded) has an error of about 2^-52 |offset /
    scale|: at offset 3 * 2^58, scale 60 that is several steps, and its
    corrected estimate misses the exact result for many powers."""
    off, sc, mode = 3.0 * 2.0 ** 58, 60.0, 0
    thr = _thresholds(oracle, off, sc, mode)
    reach = thr[1:256][thr[1:256] < INT64_MAX]
    assert reach.size > 100
    P = (reach[:, None] + np.arange(-40, 41, 8)[None, :]).reshape(-1)
    P = P[P >= 0]
    want = oracle.cb_tbtf(P.astype(np.float64).reshape(1, 1, -1), np.float32([off]), np.float32([sc]), 1,
                          mode).reshape(-1)
    np.testing.assert_array_equal(_quantise_thr(P, off, sc, thr, mode), want)
    inv = 1.0 / np.float64(np.float32(sc))
    c = -np.float64(np.float32(off)) * inv
    y_fma = P.astype(np.float64) * inv + c      # one rounding more than an fma: errors of the same size
    wrong = _settle(P, np.trunc(np.clip(y_fma, -128, 127)), thr)
    assert np.count_nonzero(wrong != want) > 100
