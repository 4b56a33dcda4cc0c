"""Coherent beams straight from TAFTP (fbf_cb_beamform_taftp) and the
incoherent beam with the weights in one launch (fbf_ib_beamform_weights).

The oracle for fbf_cb_beamform_taftp is the reference's pair as its Pipeline
calls it (Pipeline.cu:166-168): the oracle's split-transpose of the plan's
antennas, then the oracle's coherent beamformer; bit-exact.  Cases cover the
direct path (the kernel stages the TAFTP runs and transposes them in LDS:
64 or 128 beams per wave, antennas padded to a multiple of 16, subarrays at
any antenna offset, tscrunch 4 to 64, heaps of 16 to 512 samples, weight
components of -128 on the exact path) and the fallback through the plan's
FTPA buffer (geometries the direct path does not take: fscrunch above 1,
32 beams per wave, the other kernels, heaps that are not a power of two).
"""
import numpy as np
import pytest

from conftest import gaussian_int8, unit_weights

pytestmark = pytest.mark.gpu


def _dev(capi, arr):
    return capi.DeviceBuffer.from_array(arr)


def _levels(oracle, A, ts, fs, F):
    cbo, cbs, _, _ = oracle.power_levels(16.0, 24.0, A, ts, fs, 64, 16, 1)
    Fo = F // fs
    return np.full(Fo, cbo, np.float32), np.full(Fo, cbs, np.float32)


def _check_taftp(capi, oracle, Atot, A, a0, F, nhg, B, ts, fs, S, variant, direct, seed, weights=None):
    T = nhg * S
    taftp = gaussian_int8(seed, nhg * Atot * F * S * 4, 16.0)
    w = unit_weights(seed + 1, F * B * A) if weights is None else weights
    off, sc = _levels(oracle, A, ts, fs, F)
    plan = capi.CbPlan(A, B, F, ts, fs, off, sc, nsamples_per_heap_out=T // ts)
    plan.set_variant(variant)
    assert plan.taftp_direct(Atot, T, S) == direct
    ftpa = oracle.split_transpose(taftp, nhg, Atot, A, a0, F, nsamples_per_heap=S)
    ref = oracle.cb_tbtf(oracle.cb_power(ftpa, w, F, T, A, B, ts, fs), off, sc, T // ts)
    d_t, d_w = _dev(capi, taftp), _dev(capi, w)
    d_o = capi.DeviceBuffer((F // fs) * (T // ts) * B)
    d_o.fill(0x5A)
    plan.beamform_taftp(d_t, d_w, d_o, nantennas_total=Atot, start_antenna=a0, nsamples_per_heap=S)
    np.testing.assert_array_equal(d_o.download(), ref)
    return plan, d_t, d_w, d_o


@pytest.mark.parametrize("Atot,A,a0,F,nhg,B,ts,fs,S,variant", [
    (64, 64, 0, 8, 2, 512, 16, 1, 256, 0),     # the bench geometry (channels and heaps cut)
    (64, 64, 0, 4, 2, 512, 16, 1, 256, 11),    # 128 beams per wave
    (64, 40, 0, 8, 2, 96, 16, 1, 256, 11),     # 40 antennas: rows 40..47 padded
    (64, 64, 0, 4, 2, 256, 4, 1, 256, 9),      # tscrunch 4 (four windows a tile)
    (64, 32, 32, 4, 2, 128, 4, 1, 256, 9),     # upper half, tscrunch 4
    (64, 16, 40, 8, 2, 64, 16, 1, 256, 9),     # subarray at antenna 40
    (64, 8, 56, 4, 2, 64, 8, 1, 256, 9),       # last 8 antennas, one K step
    (128, 64, 64, 2, 2, 256, 32, 1, 256, 9),   # upper half of a 128-antenna block, tscrunch 32
    (64, 48, 8, 4, 2, 200, 64, 1, 256, 11),    # tscrunch 64 (four tiles a window), ragged beams
    (64, 64, 0, 4, 4, 512, 16, 1, 128, 9),     # 128-sample heaps
    (64, 64, 0, 2, 1, 256, 16, 1, 512, 9),     # 512-sample heaps
    (32, 24, 4, 4, 16, 64, 16, 1, 16, 11),     # 16-sample heaps: one tile a heap
    (64, 64, 0, 4, 3, 512, 8, 1, 256, 11),     # three heap groups, tscrunch 8
])
def test_cb_taftp_direct_bit_exact(gpu, oracle, Atot, A, a0, F, nhg, B, ts, fs, S, variant):
    _check_taftp(gpu, oracle, Atot, A, a0, F, nhg, B, ts, fs, S, variant, True, 2000 + A + a0 + S)


@pytest.mark.parametrize("Atot,A,a0,F,nhg,B,ts,fs,S,variant", [
    (64, 40, 0, 8, 2, 96, 4, 2, 256, 0),       # 40 antennas: the automatic pick is the wave-private kernel
    (128, 128, 0, 2, 1, 32, 8, 1, 256, 0),     # 128 antennas: 32 beams per wave
    (64, 20, 3, 4, 2, 64, 16, 1, 256, 0),      # 20 antennas (the LDS kernel)
    (64, 64, 0, 4, 2, 512, 16, 1, 96, 9),      # 96-sample heaps (not a power of two)
    (64, 64, 0, 4, 2, 256, 16, 1, 256, 5),     # the LDS variant chosen explicitly
    (64, 32, 16, 8, 2, 128, 8, 4, 256, 9),     # fscrunch 4
    (64, 64, 0, 4, 2, 512, 32, 2, 256, 8),     # fscrunch 2, 32 beams per wave
])
def test_cb_taftp_fallback_bit_exact(gpu, oracle, Atot, A, a0, F, nhg, B, ts, fs, S, variant):
    _check_taftp(gpu, oracle, Atot, A, a0, F, nhg, B, ts, fs, S, variant, False, 2100 + A + a0 + S)


@pytest.mark.parametrize("variant,Atot,A,a0", [(9, 128, 64, 32), (11, 64, 48, 8)])
def test_cb_taftp_minus128_weights_exact(gpu, oracle, variant, Atot, A, a0):
    """Weight components of -128 on the direct path: the work-groups that
    hold one compute their outputs on the exact int32 path, which then reads
    the TAFTP block itself (antenna stride F * S * 4 bytes)."""
    F, nhg, B, ts, fs, S = 4, 2, 256, 16, 1, 256
    w = unit_weights(2201, F * B * A).reshape(F, B, A, 2).copy()
    w[0, 3, 5, 1] = -128
    w[F - 1, B - 7, 0, 1] = -128
    w[F - 1, B - 7, 1, 0] = -128
    w[1, B // 2, A - 1, :] = -128
    _check_taftp(gpu, oracle, Atot, A, a0, F, nhg, B, ts, fs, S, variant, True, 2200, weights=w.reshape(-1))


def test_cb_taftp_direct_needs_block_below_4gib(gpu, oracle):
    """The direct path addresses the block with 32-bit tile offsets, so a
    block of 4 GiB or more takes the split-transpose path (a query only: no
    such block is allocated)."""
    capi = gpu
    A, F, B, ts, S = 64, 64, 512, 16, 256
    off, sc = _levels(oracle, A, ts, 1, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc)
    hg = 4 * A * F * S                    # bytes per heap group
    below = ((1 << 32) - 1) // hg          # heap groups of the largest block below 4 GiB
    assert plan.taftp_direct(A, below * S, S)
    assert not plan.taftp_direct(A, (below + 1) * S, S)


def test_cb_taftp_matches_ftpa_path_and_repeats(gpu, oracle):
    """The direct path equals split-transpose + fbf_cb_beamform on the GPU
    at a larger block (64 channels, 4 heap groups, 512 beams), on two calls
    in a row (the second with other voltages), and a forced fallback (the
    LDS variant) equals both."""
    capi = gpu
    Atot, A, F, nhg, B, ts, S = 64, 64, 64, 4, 512, 16, 256
    T = nhg * S
    off, sc = _levels(oracle, A, ts, 1, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts)
    assert plan.taftp_direct(Atot, T, S)
    w = unit_weights(2301, F * B * A)
    d_w = _dev(capi, w)
    d_f = capi.DeviceBuffer(F * T * 2 * A * 2)
    d_a = capi.DeviceBuffer(F * (T // ts) * B)
    d_b = capi.DeviceBuffer(F * (T // ts) * B)
    for seed in (2302, 2303):
        d_t = _dev(capi, gaussian_int8(seed, nhg * Atot * F * S * 4, 16.0))
        capi.split_transpose(d_t, d_f, Atot, A, 0, F)
        plan.beamform(d_f, d_w, d_a)
        d_b.fill(0x5A)
        plan.beamform_taftp(d_t, d_w, d_b, nantennas_total=Atot)
        np.testing.assert_array_equal(d_b.download(), d_a.download())
    plan.set_variant(5)
    assert not plan.taftp_direct(Atot, T, S)
    d_b.fill(0x5A)
    plan.beamform_taftp(d_t, d_w, d_b, nantennas_total=Atot)
    np.testing.assert_array_equal(d_b.download(), d_a.download())


def test_cb_taftp_rejects_bad_arguments(gpu):
    capi = gpu
    A, B, F, ts, S = 16, 32, 4, 16, 256
    off, sc = np.full(F, 1.0, np.float32), np.full(F, 1.0, np.float32)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=S // ts)
    d_t = capi.DeviceBuffer(32 * F * S * 4)       # one heap group of 32 antennas
    d_w = capi.DeviceBuffer(F * B * A * 2)
    d_o = capi.DeviceBuffer(F * (S // ts) * B)
    with pytest.raises(capi.FbfError):            # antennas past the block's
        plan.beamform_taftp(d_t, d_w, d_o, nantennas_total=32, start_antenna=20)
    with pytest.raises(capi.FbfError):            # not a whole number of heap groups
        plan.beamform_taftp(d_t, d_w, d_o, nantennas_total=48)
    with pytest.raises(capi.FbfError):            # weights size
        plan.beamform_taftp(d_t, capi.DeviceBuffer(F * B * A * 2 - 2), d_o, nantennas_total=32)
    with pytest.raises(capi.FbfError):            # output size
        plan.beamform_taftp(d_t, d_w, capi.DeviceBuffer(F * (S // ts) * B + 1), nantennas_total=32)
    plan.set_variant(5)                           # the fallback checks sizes before it launches
    with pytest.raises(capi.FbfError):
        plan.beamform_taftp(d_t, capi.DeviceBuffer(F * B * A * 2 - 2), d_o, nantennas_total=32)
    with pytest.raises(capi.FbfError):
        plan.beamform_taftp(d_t, d_w, capi.DeviceBuffer(F * (S // ts) * B + 1), nantennas_total=32)
    plan.beamform_taftp(d_t, d_w, d_o, nantennas_total=32, start_antenna=16)


@pytest.mark.parametrize("A,F,nhg,ts,fs,B", [(64, 64, 4, 16, 1, 512), (40, 8, 2, 32, 2, 96), (8, 16, 3, 1, 1, 7),
                                             (24, 4, 1, 4, 4, 33)])
def test_ib_beamform_weights_fused(gpu, oracle, A, F, nhg, ts, fs, B):
    """The incoherent beam and the weights in one launch equal
    fbf_ib_beamform and fbf_weights_generate byte for byte, and the IB
    output equals the oracle's."""
    capi = gpu
    taftp = gaussian_int8(2400 + A, nhg * A * F * 256 * 4, 16.0)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, 16, 1, A, ts, fs)
    Fo = F // fs
    ioff, isc = np.full(Fo, ibo, np.float32), np.full(Fo, ibs, np.float32)
    ib = capi.IbPlan(A, F, ts, fs, ioff, isc)
    freqs = capi.channel_frequencies_f64(1284e6, 856e6, F)
    wp = capi.WeightsPlan(A, B, F, freqs)
    delays = np.zeros((B * A, 2), np.float32)
    delays[:, 0] = capi.synth_uniform_f32(2401, -1e-9, 1e-9, B * A)
    delays[:, 1] = capi.synth_uniform_f32(2402, -1e-6, 1e-6, B * A)
    d_t, d_d = _dev(capi, taftp), _dev(capi, delays.reshape(-1))
    d_ib, d_ib2 = capi.DeviceBuffer(ib.output_bytes(d_t.nbytes)), capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    d_w, d_w2 = capi.DeviceBuffer(F * B * A * 2), capi.DeviceBuffer(F * B * A * 2)
    ib.beamform(d_t, d_ib)
    wp.generate(d_d, 12.5, d_w)
    d_ib2.fill(0x5A)
    d_w2.fill(0x5A)
    ib.beamform_weights(d_t, d_ib2, wp, d_d, 12.5, d_w2)
    np.testing.assert_array_equal(d_ib2.download(), d_ib.download())
    np.testing.assert_array_equal(d_w2.download(), d_w.download())
    ref = oracle.ib_tf(oracle.ib_power(taftp, F, ts, fs, nhg, A), ioff, isc)
    np.testing.assert_array_equal(d_ib2.download(), ref.reshape(-1))


def test_ib_beamform_weights_rejects_channel_mismatch(gpu):
    capi = gpu
    A, F, B = 8, 16, 8
    ib = capi.IbPlan(A, F, 16, 1, np.zeros(F, np.float32), np.ones(F, np.float32))
    wp = capi.WeightsPlan(A, B, F // 2, capi.channel_frequencies_f64(1284e6, 856e6, F // 2))
    d_t = capi.DeviceBuffer(A * F * 256 * 4)
    d_ib = capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    d_d = capi.DeviceBuffer(B * A * 8)
    d_w = capi.DeviceBuffer((F // 2) * B * A * 2)
    with pytest.raises(capi.FbfError):
        ib.beamform_weights(d_t, d_ib, wp, d_d, 0.0, d_w)


def _check_cb_ib(capi, oracle, A, F, nhg, B, ts, ib_ts, ib_fs, S, variant, fused, seed, weights=None):
    """fbf_cb_ib_beamform_taftp against the oracle: the coherent beams as in
    _check_taftp, the incoherent beam as the oracle's ib_power + ib_tf over
    all antennas (IncoherentBeamformer.cu:10-69)."""
    T = nhg * S
    taftp = gaussian_int8(seed, nhg * A * F * S * 4, 16.0)
    w = unit_weights(seed + 1, F * B * A) if weights is None else weights
    off, sc = _levels(oracle, A, ts, 1, F)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, ts, 1, A, ib_ts, ib_fs)
    Fo = F // ib_fs
    ioff, isc = np.full(Fo, ibo, np.float32), np.full(Fo, ibs, np.float32)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts)
    if variant:
        plan.set_variant(variant)
    ib = capi.IbPlan(A, F, ib_ts, ib_fs, ioff, isc, nsamples_per_heap=S)
    assert plan.ib_fused(ib, T, S) == fused
    ftpa = oracle.split_transpose(taftp, nhg, A, A, 0, F, nsamples_per_heap=S)
    ref = oracle.cb_tbtf(oracle.cb_power(ftpa, w, F, T, A, B, ts, 1), off, sc, T // ts)
    ib_ref = oracle.ib_tf(oracle.ib_power(taftp, F, ib_ts, ib_fs, nhg, A, nsamples_per_heap=S), ioff, isc)
    d_t, d_w = _dev(capi, taftp), _dev(capi, w)
    d_o = capi.DeviceBuffer(F * (T // ts) * B)
    d_tf = capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    for _ in range(2):   # a second call finds the partials buffer in use
        d_o.fill(0x5A)
        d_tf.fill(0x5A)
        plan.ib_beamform_taftp(ib, d_t, d_w, d_o, d_tf, nsamples_per_heap=S)
        np.testing.assert_array_equal(d_o.download(), ref)
        np.testing.assert_array_equal(d_tf.download(), ib_ref.reshape(-1))


@pytest.mark.parametrize("A,F,nhg,B,ts,ib_ts,ib_fs,S,variant", [
    (64, 8, 2, 512, 16, 16, 1, 256, 0),     # the bench geometry (channels and heaps cut)
    (64, 4, 2, 256, 4, 1, 1, 256, 0),       # IB per sample, CB tscrunch 4
    (64, 4, 3, 512, 8, 256, 2, 256, 0),     # IB windows of a whole heap, fscrunch 2
    (64, 8, 2, 1024, 32, 8, 4, 256, 0),     # two beam groups (only the first leaves partials), IB fscrunch 4
    (40, 8, 2, 200, 16, 16, 1, 256, 11),    # antennas 40..47 padded; waves without beams
    (48, 4, 2, 128, 64, 32, 1, 256, 11),    # three K slices: the fourth wave's slot holds zeros
    (16, 4, 4, 128, 16, 16, 1, 256, 11),    # one K slice
])
def test_cb_ib_fused_bit_exact(gpu, oracle, A, F, nhg, B, ts, ib_ts, ib_fs, S, variant):
    _check_cb_ib(gpu, oracle, A, F, nhg, B, ts, ib_ts, ib_fs, S, variant, True, 2500 + A + B + ib_ts)


@pytest.mark.parametrize("A,F,nhg,B,ts,ib_ts,ib_fs,S,variant", [
    (64, 4, 2, 256, 16, 16, 1, 256, 9),     # 64 beams per wave: the two beamformers in turn
    (40, 8, 2, 96, 16, 16, 2, 256, 0),      # the automatic pick at 40 antennas
])
def test_cb_ib_unfused_bit_exact(gpu, oracle, A, F, nhg, B, ts, ib_ts, ib_fs, S, variant):
    _check_cb_ib(gpu, oracle, A, F, nhg, B, ts, ib_ts, ib_fs, S, variant, False, 2600 + A + B + ib_ts)


def test_cb_ib_fused_minus128_weights_exact(gpu, oracle):
    """Work-groups with a weight component of -128 take the exact path and
    leave the incoherent-beam partials from the TAFTP block themselves."""
    A, F, nhg, B = 64, 4, 2, 512
    w = unit_weights(2701, F * B * A).reshape(F, B, A, 2).copy()
    w[0, 3, 5, 1] = -128
    w[F - 1, B - 7, 0, 1] = -128
    w[2, 100, :, :] = -128
    _check_cb_ib(gpu, oracle, A, F, nhg, B, 16, 16, 1, 256, 0, True, 2700, weights=w.reshape(-1))


@pytest.mark.parametrize("A,B,variant,alias", [(64, 512, 0, True), (64, 256, 0, False), (64, 256, 9, True)])
def test_cb_ib_next_weights(gpu, oracle, A, B, variant, alias):
    """fbf_cb_ib_beamform_taftp_next equals fbf_cb_ib_beamform_taftp followed
    by fbf_weights_generate at the next epoch: beams from this block's
    weights, then the next block's weights, also when they overwrite this
    block's (alias) and on the unfused path (variant 9)."""
    capi = gpu
    F, nhg, ts, S = 8, 2, 16, 256
    T = nhg * S
    taftp = gaussian_int8(2800 + B, nhg * A * F * S * 4, 16.0)
    off, sc = _levels(oracle, A, ts, 1, F)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, ts, 1, A, 16, 1)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts)
    if variant:
        plan.set_variant(variant)
    ib = capi.IbPlan(A, F, 16, 1, np.full(F, ibo, np.float32), np.full(F, ibs, np.float32))
    assert plan.ib_fused(ib, T) == (variant == 0)
    freqs = capi.channel_frequencies_f64(1284e6, 856e6, F)
    wp = capi.WeightsPlan(A, B, F, freqs)
    delays = np.zeros((B * A, 2), np.float32)
    delays[:, 0] = capi.synth_uniform_f32(2801, -1e-9, 1e-9, B * A)
    delays[:, 1] = capi.synth_uniform_f32(2802, -1e-6, 1e-6, B * A)
    d_t, d_d = _dev(capi, taftp), _dev(capi, delays.reshape(-1))
    d_w = capi.DeviceBuffer(F * B * A * 2)
    wp.generate(d_d, 3.0, d_w)
    w_now = d_w.download()
    d_wn = capi.DeviceBuffer(F * B * A * 2)
    wp.generate(d_d, 3.0 + 0.0392, d_wn)
    w_next = d_wn.download()
    d_cb, d_tf = capi.DeviceBuffer(F * (T // ts) * B), capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    d_cb2, d_tf2 = capi.DeviceBuffer(F * (T // ts) * B), capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    plan.ib_beamform_taftp(ib, d_t, d_w, d_cb, d_tf)
    d_out_w = d_w if alias else capi.DeviceBuffer(F * B * A * 2)
    d_cb2.fill(0x5A)
    d_tf2.fill(0x5A)
    plan.ib_beamform_taftp_next(ib, d_t, d_w, d_cb2, d_tf2, wp, d_d, 3.0 + 0.0392, d_out_w)
    np.testing.assert_array_equal(d_cb2.download(), d_cb.download())
    np.testing.assert_array_equal(d_tf2.download(), d_tf.download())
    np.testing.assert_array_equal(d_out_w.download(), w_next)
    if not alias:
        np.testing.assert_array_equal(d_w.download(), w_now)
    ftpa = oracle.split_transpose(taftp, nhg, A, A, 0, F)
    ref = oracle.cb_tbtf(oracle.cb_power(ftpa, w_now, F, T, A, B, ts, 1), off, sc, T // ts)
    np.testing.assert_array_equal(d_cb2.download(), ref)


def test_cb_plan_timing_events(gpu, oracle):
    """fbf_cb_plan_set_timing: the plan records the events around the
    coherent kernel alone, inside the call's own span, on every later call
    until unset; a half-set pair is refused."""
    capi = gpu
    A, F, nhg, B, ts, S = 64, 8, 2, 512, 16, 256
    T = nhg * S
    off, sc = _levels(oracle, A, ts, 1, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, ts, 1, A, 16, 1)
    ib = capi.IbPlan(A, F, 16, 1, np.full(F, ibo, np.float32), np.full(F, ibs, np.float32))
    d_t = _dev(capi, gaussian_int8(2900, nhg * A * F * S * 4, 16.0))
    d_w = _dev(capi, unit_weights(2901, F * B * A))
    d_cb, d_tf = capi.DeviceBuffer(F * (T // ts) * B), capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    s = capi.Stream()
    c0, c1, k0, k1 = capi.Event(), capi.Event(), capi.Event(), capi.Event()
    plan.set_timing(k0, k1)
    c0.record(s)
    plan.ib_beamform_taftp(ib, d_t, d_w, d_cb, d_tf, stream=s)
    c1.record(s)
    s.sync()
    kern, call, before = k0.elapsed_ms(k1), c0.elapsed_ms(c1), c0.elapsed_ms(k0)
    assert 0 < kern <= call and 0 <= before <= call - kern
    plan.set_timing(None, None)
    plan.ib_beamform_taftp(ib, d_t, d_w, d_cb, d_tf, stream=s)
    s.sync()
    assert k0.elapsed_ms(k1) == kern   # not recorded again
    with pytest.raises(capi.FbfError):
        capi.call("fbf_cb_plan_set_timing", plan.handle, k0.handle, None)


def test_cb_ib_next_weights_rejects_other_geometry(gpu, oracle):
    """A weights plan of other channels with the same byte count (half the
    channels, twice the beams) is refused before anything is launched.  (A
    plan of another antennas x beams split with the same product gives the
    same bytes: the generator works per (beam, antenna) index.)"""
    capi = gpu
    A, B, F, ts, S = 64, 256, 4, 16, 256
    off, sc = _levels(oracle, A, ts, 1, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=S // ts)
    ib = capi.IbPlan(A, F, 16, 1, np.zeros(F, np.float32), np.ones(F, np.float32))
    freqs = capi.channel_frequencies_f64(1284e6, 856e6, F)
    wp_other = capi.WeightsPlan(A, 2 * B, F // 2, freqs[: F // 2])   # same bytes, other channels
    d_t = capi.DeviceBuffer(A * F * S * 4)
    d_w = capi.DeviceBuffer(F * B * A * 2)
    d_d = capi.DeviceBuffer(2 * B * A * 8)
    d_cb, d_tf = capi.DeviceBuffer(F * (S // ts) * B), capi.DeviceBuffer(ib.output_bytes(d_t.nbytes))
    with pytest.raises(capi.FbfError):
        plan.ib_beamform_taftp_next(ib, d_t, d_w, d_cb, d_tf, wp_other, d_d, 1.0, d_w)
