"""Parity of the HIP kernels (through the C ABI) against the CPU oracle.

Bit-exact for split-transpose, int32 beam sums, exact power and the int8
outputs (both the kernels and the oracle evaluate the quantiser on the exact
power in IEEE double).  Weights: FP32 sin/cos on the GPU vs the FP64 oracle,
tolerance stated in the tests.
"""
import numpy as np
import pytest

from conftest import gaussian_int8, unit_weights
import refmath

pytestmark = pytest.mark.gpu


def _dev(capi, arr):
    return capi.DeviceBuffer.from_array(arr)


# ---------------------------------------------------------------------------
# split transpose
# ---------------------------------------------------------------------------
def _cycling_prime(n):
    i = np.arange(n)
    t = np.empty((n, 2), np.int8)
    t[:, 0] = (i % 113).astype(np.int8)
    t[:, 1] = (i % 107).astype(np.int8)
    return t.reshape(-1)


@pytest.mark.parametrize("A,used,a0,F,nhg", [
    (64, 44, 0, 64, 12),      # reference geometry (fbfuse_constants.hpp:4-13), SplitTransposeTester.cu:96-113
    (64, 64, 0, 64, 4),
    (64, 40, 0, 64, 2), (64, 16, 40, 64, 2), (64, 8, 56, 64, 2),   # 40/16/8 split
    (64, 6, 3, 8, 2), (16, 3, 5, 4, 2), (128, 100, 20, 4, 1),
])
def test_split_transpose_bit_exact(gpu, oracle, A, used, a0, F, nhg):
    capi = gpu
    n = nhg * A * F * 256 * 2
    taftp = _cycling_prime(n)
    d_in = _dev(capi, taftp)
    d_out = capi.DeviceBuffer(n // A * used * 2)
    capi.split_transpose(d_in, d_out, A, used, a0, F)
    got = d_out.download()
    np.testing.assert_array_equal(got, oracle.split_transpose(taftp, nhg, A, used, a0, F))


# ---------------------------------------------------------------------------
# coherent beamformer
# ---------------------------------------------------------------------------
def _levels(oracle, A, ts, fs, F, level_in=16.0, level_out=24.0):
    cbo, cbs, _, _ = oracle.power_levels(level_in, level_out, A, ts, fs, 64, 16, 1)
    Fo = F // fs
    return np.full(Fo, cbo, np.float32), np.full(Fo, cbs, np.float32)


CB_SUMS_CASES = [
    # F, T, A, B
    (16, 2048, 16, 32),     # config 1 geometry
    (4, 512, 44, 64),       # 44 antennas: K = 88, partial K slice
    (2, 256, 64, 512),      # config 2 antennas/beams
    (2, 256, 40, 768), (2, 256, 16, 256), (2, 256, 8, 64),   # config 3 subarrays
    (1, 256, 128, 24), (3, 128, 4, 8), (1, 64, 32, 40),
]


@pytest.mark.parametrize("F,T,A,B", CB_SUMS_CASES)
def test_cb_int32_sums_bit_exact(gpu, oracle, F, T, A, B):
    capi = gpu
    ftpa = gaussian_int8(100 + A, F * T * 2 * A * 2, 16.0)
    w = unit_weights(200 + B, F * B * A)
    off, sc = _levels(oracle, A, 16, 1, F)
    plan = capi.CbPlan(A, B, F, 16 if T % 16 == 0 else 1, 1, off, sc)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_s = capi.DeviceBuffer(F * T * 2 * B * 2 * 4)
    plan.beam_sums(d_v, d_w, d_s)
    got = d_s.download(np.int32)
    ref = oracle.cb_sums(ftpa, w, F, T, A, B)
    np.testing.assert_array_equal(got, ref.reshape(-1))


@pytest.mark.parametrize("F,T,A,B,ts,fs", [
    (16, 2048, 16, 32, 16, 1), (4, 1024, 64, 512, 16, 1), (8, 512, 44, 128, 16, 4),
    (4, 1024, 40, 768, 64, 1), (4, 1024, 16, 256, 4, 2), (4, 512, 8, 64, 1, 1), (2, 256, 128, 16, 8, 2),
])
def test_cb_power_exact(gpu, oracle, F, T, A, B, ts, fs):
    capi = gpu
    ftpa = gaussian_int8(300 + A, F * T * 2 * A * 2, 16.0)
    w = unit_weights(400 + B, F * B * A)
    off, sc = _levels(oracle, A, ts, fs, F)
    plan = capi.CbPlan(A, B, F, ts, fs, off, sc)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_p = capi.DeviceBuffer((F // fs) * (T // ts) * B * 8)
    plan.beam_power(d_v, d_w, d_p)
    got = d_p.download(np.float64).reshape(F // fs, T // ts, B)
    np.testing.assert_array_equal(got, oracle.cb_power(ftpa, w, F, T, A, B, ts, fs))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("F,T,A,B,ts,fs,ohs", [
    (16, 2048, 16, 32, 16, 1, 128),         # config 1 (2048-byte output heaps)
    (8, 2048, 64, 512, 16, 1, 0),           # config 2 geometry, 8 channels, default heap (1024 samples)
    (8, 1024, 44, 64, 16, 4, 64),
    (4, 4096, 40, 768, 64, 1, 0), (4, 1024, 8, 64, 4, 1, 0),
    # TBTF reorder edge tiles: channels not a multiple of 4, more than 64
    # channels, partial beam tiles, heaps that are not a multiple of the tile
    (6, 768, 8, 40, 2, 1, 24), (132, 64, 4, 8, 2, 1, 8), (10, 512, 12, 72, 4, 2, 0), (20, 1200, 16, 24, 3, 1, 50),
])
def test_cb_int8_tbtf_bit_exact(gpu, oracle, mode, F, T, A, B, ts, fs, ohs):
    capi = gpu
    ftpa = gaussian_int8(500 + A, F * T * 2 * A * 2, 16.0)
    w = unit_weights(600 + B, F * B * A)
    off, sc = _levels(oracle, A, ts, fs, F)
    Fo, To = F // fs, T // ts
    ohs_eff = ohs if ohs else 8192 // Fo
    if To % ohs_eff:
        ohs_eff = To
        ohs = To
    plan = capi.CbPlan(A, B, F, ts, fs, off, sc, nsamples_per_heap_out=ohs, quant_mode=mode)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    nout = plan.output_bytes(ftpa.nbytes)
    assert nout == To * B * Fo
    d_o = capi.DeviceBuffer(nout)
    plan.beamform(d_v, d_w, d_o)
    got = d_o.download()
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, fs)
    ref = oracle.cb_tbtf(p, off, sc, ohs_eff, mode)
    np.testing.assert_array_equal(got, ref)
    # outputs are not degenerate: the levels put the power inside the int8 range
    assert got.astype(float).std() > 3.0 and np.mean(np.abs(got.astype(int)) == 127) < 0.01


def test_cb_matches_reference_f32_within_reference_tolerance(gpu, oracle):
    """Reference test geometry (44 antennas, ts 16, fs 4, truncating quantiser,
    input/output level 32, unclipped N(0, 32) voltages; CoherentBeamformerTester.cu:148-185)
    at 8 channels and 256 beams: every output within +-1 of the reference's FP32 model."""
    capi = gpu
    F, T, A, B, ts, fs = 8, 1024, 44, 256, 16, 4
    ftpa = gaussian_int8(700, F * T * 2 * A * 2, 32.0, clip=False)
    w = unit_weights(701, F * B * A)
    cbo, cbs, _, _ = oracle.power_levels(32.0, 32.0, A, ts, fs, 64, 16, 2)
    Fo, To = F // fs, T // ts
    plan = capi.CbPlan(A, B, F, ts, fs, np.full(Fo, cbo, np.float32), np.full(Fo, cbs, np.float32),
                       nsamples_per_heap_out=To, quant_mode=1)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_o = capi.DeviceBuffer(To * B * Fo)
    plan.beamform(d_v, d_w, d_o)
    got = d_o.download()
    ref = oracle.cb_reference_f32(ftpa, w, F, ts, fs, T, B, A, cbs, cbo, To)
    d = np.abs(got.astype(int) - ref.astype(int))
    print(f"GPU vs FP32 reference model: {(d > 0).sum()} of {d.size} outputs differ ({(d > 0).mean():.2e})")
    assert d.max() <= 1
    # 1 of 32 768 outputs differs for this seed: the FP32 power rounding of
    # the reference model crosses one quantiser step (the GPU output equals
    # the exact oracle bit for bit)
    assert (d > 0).mean() <= 1e-4


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11])
@pytest.mark.parametrize("weights", ["sparse", "full_range"])
@pytest.mark.parametrize("F,T,A,B,ts,fs", [(4, 512, 64, 256, 16, 1), (4, 512, 40, 96, 4, 2),
                                            (2, 1024, 16, 64, 64, 1)])
def test_cb_minus128_weights_exact(gpu, oracle, variant, weights, F, T, A, B, ts, fs):
    """Weight components of -128: the int32 sums follow the reference's exact
    arithmetic r = wr vr - wi vi (CoherentBeamformer.cu:150-155), so
    wi = -128 contributes +128 vi, and the power and int8 outputs follow.
    'sparse': unit-phase weights with a few -128 components (wi alone, wr
    alone, both, a whole beam), so one launch has waves on the MFMA path and
    waves on the exact path; 'full_range': uniform int8 weights over
    [-128, 127], where nearly every wave takes the exact path.  Every kernel
    variant, sums, power and int8 against the oracle, and the sums also
    against the numpy complex product."""
    capi = gpu
    ftpa = gaussian_int8(810 + A, F * T * 2 * A * 2, 30.0, clip=False)
    if weights == "sparse":
        w = unit_weights(811 + B, F * B * A).reshape(F, B, A, 2).copy()
        w[0, 3, 5, 1] = -128
        w[F - 1, B - 7, 0, 1] = -128
        w[F - 1, B - 7, 1, 0] = -128
        w[1, B // 2, A - 1, :] = -128
        w[1, 9, :, 1] = -128
    else:
        w = np.random.default_rng(812 + A).integers(-128, 128, F * B * A * 2).astype(np.int8)
    w = w.reshape(-1)
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, fs)
    # levels from the power itself (median, spread / 40): few outputs
    # saturate and one quantiser step is a fraction of a percent of the power
    off = np.median(p.reshape(F // fs, -1), axis=1).astype(np.float32)
    sc = (p.reshape(F // fs, -1).std(axis=1) / 40).astype(np.float32)
    ref_q = oracle.cb_tbtf(p, off, sc, T // ts)
    # the inputs tell the exact arithmetic from a saturated -(-128) = 127 at int8
    wsat = w.reshape(F, B, A, 2).astype(np.int64)
    v = ftpa.reshape(F, T, 2, A, 2).astype(np.int64)
    nwi = np.where(wsat[..., 1] == -128, 127, -wsat[..., 1])
    re = np.einsum("ftpa,fba->ftpb", v[..., 0], wsat[..., 0]) + np.einsum("ftpa,fba->ftpb", v[..., 1], nwi)
    im = np.einsum("ftpa,fba->ftpb", v[..., 1], wsat[..., 0]) + np.einsum("ftpa,fba->ftpb", v[..., 0], wsat[..., 1])
    p_sat = (re ** 2 + im ** 2).sum(axis=2).reshape(F // fs, fs, T // ts, ts, B).sum(axis=(1, 3)).astype(np.float64)
    assert np.count_nonzero(oracle.cb_tbtf(p_sat, off, sc, T // ts) != ref_q) > 0
    plan = capi.CbPlan(A, B, F, ts, fs, off, sc, nsamples_per_heap_out=T // ts)
    plan.set_variant(variant)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    ref_sums = oracle.cb_sums(ftpa, w, F, T, A, B)
    np.testing.assert_array_equal(ref_sums, refmath.cb_sums_np(ftpa, w, F, T, A, B))
    d_s = capi.DeviceBuffer(F * T * 2 * B * 2 * 4)
    plan.beam_sums(d_v, d_w, d_s)
    np.testing.assert_array_equal(d_s.download(np.int32), ref_sums.reshape(-1))
    d_p = capi.DeviceBuffer((F // fs) * (T // ts) * B * 8)
    plan.beam_power(d_v, d_w, d_p)
    np.testing.assert_array_equal(d_p.download(np.float64).reshape(F // fs, T // ts, B), p)
    d_o = capi.DeviceBuffer((F // fs) * (T // ts) * B)
    plan.beamform(d_v, d_w, d_o)
    np.testing.assert_array_equal(d_o.download(), ref_q)


def test_cb_unit_weights_closed_form(gpu):
    """w = (127, 0): beam = 127 * sum_a v (computed here with numpy)."""
    capi = gpu
    F, T, A, B = 2, 256, 64, 64
    ftpa = gaussian_int8(900, F * T * 2 * A * 2, 16.0)
    w = np.zeros(F * B * A * 2, np.int8)
    w[0::2] = 127
    plan = capi.CbPlan(A, B, F, 16, 1, np.ones(F, np.float32), np.ones(F, np.float32))
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_s = capi.DeviceBuffer(F * T * 2 * B * 2 * 4)
    plan.beam_sums(d_v, d_w, d_s)
    s = d_s.download(np.int32).reshape(F, T, 2, B, 2)
    v = ftpa.reshape(F, T, 2, A, 2).astype(np.int64).sum(axis=3)
    for b in (0, 17, 63):
        np.testing.assert_array_equal(s[:, :, :, b, :], 127 * v)


def test_cb_rejects_bad_sizes(gpu):
    capi = gpu
    plan = capi.CbPlan(64, 512, 4, 16, 1, np.ones(4, np.float32), np.ones(4, np.float32))
    d_v = capi.DeviceBuffer(4 * 256 * 2 * 64 * 2)
    d_w = capi.DeviceBuffer(4 * 512 * 64 * 2 - 2)
    d_o = capi.DeviceBuffer(16 * 512 * 4)
    with pytest.raises(capi.FbfError) as e:
        plan.beamform(d_v, d_w, d_o)
    assert e.value.code == capi.FBF_ERR_SIZE


# ---------------------------------------------------------------------------
# incoherent beamformer
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("A,F,nhg,ts,fs", [
    (64, 64, 4, 16, 1), (64, 64, 2, 16, 2), (64, 64, 2, 1, 1), (64, 64, 2, 2, 4),
    (16, 16, 8, 16, 1), (40, 32, 2, 64, 8), (8, 8, 3, 4, 1), (64, 64, 1, 256, 16),
])
def test_ib_bit_exact(gpu, oracle, A, F, nhg, ts, fs):
    capi = gpu
    taftp = gaussian_int8(1000 + A + ts, nhg * A * F * 256 * 4, 16.0)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, 16, 1, A, ts, fs)
    Fo = F // fs
    off, sc = np.full(Fo, ibo, np.float32), np.full(Fo, ibs, np.float32)
    d_in = _dev(capi, taftp)
    p_ref = oracle.ib_power(taftp, F, ts, fs, nhg, A)
    for mode in (0, 1):
        plan = capi.IbPlan(A, F, ts, fs, off, sc, quant_mode=mode)
        d_p = capi.DeviceBuffer(p_ref.size * 8)
        plan.power(d_in, d_p)
        np.testing.assert_array_equal(d_p.download(np.float64).reshape(p_ref.shape), p_ref)
        d_o = capi.DeviceBuffer(plan.output_bytes(taftp.nbytes))
        plan.beamform(d_in, d_o)
        np.testing.assert_array_equal(d_o.download(), oracle.ib_tf(p_ref, off, sc, mode))


@pytest.mark.parametrize("A,F,nhg,ts", [(64, 64, 4, 16), (40, 8, 2, 32), (8, 16, 3, 1), (24, 4, 1, 4),
                                         (128, 4, 2, 16), (12, 8, 2, 8)])
def test_split_transpose_ib_fused(gpu, oracle, A, F, nhg, ts):
    """One pass over a full-antenna block gives the oracle's FTPA and TF."""
    capi = gpu
    taftp = gaussian_int8(1200 + A + ts, nhg * A * F * 256 * 4, 16.0)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, 16, 1, A, ts, 1)
    off, sc = np.full(F, ibo, np.float32), np.full(F, ibs, np.float32)
    d_in = _dev(capi, taftp)
    d_ftpa = capi.DeviceBuffer(taftp.nbytes)
    p_ref = oracle.ib_power(taftp, F, ts, 1, nhg, A)
    for mode in (0, 1):
        plan = capi.IbPlan(A, F, ts, 1, off, sc, quant_mode=mode)
        d_tf = capi.DeviceBuffer(plan.output_bytes(taftp.nbytes))
        d_ftpa.fill(0x5A)
        capi.split_transpose_ib(d_in, d_ftpa, F, plan, d_tf)
        np.testing.assert_array_equal(d_ftpa.download(), oracle.split_transpose(taftp, nhg, A, A, 0, F))
        np.testing.assert_array_equal(d_tf.download(), oracle.ib_tf(p_ref, off, sc, mode))


@pytest.mark.parametrize("A,F,nhg,ts,B", [(64, 64, 4, 16, 512), (40, 8, 2, 32, 96), (8, 16, 3, 1, 7),
                                           (24, 4, 1, 4, 300)])
def test_split_transpose_ib_weights_fused(gpu, oracle, A, F, nhg, ts, B):
    """The weights generated by the first work-groups of the transpose launch
    equal fbf_weights_generate's byte for byte, and the FTPA and TF outputs
    still equal the oracle's (beam counts that leave partial work-groups, and
    channel counts below and above one channel run)."""
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    taftp = gaussian_int8(1250 + A + ts, nhg * A * F * 256 * 4, 16.0)
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, A, 16, 1, A, ts, 1)
    off, sc = np.full(F, ibo, np.float32), np.full(F, ibs, np.float32)
    d_in = _dev(capi, taftp)
    d_ftpa = capi.DeviceBuffer(taftp.nbytes)
    plan = capi.IbPlan(A, F, ts, 1, off, sc)
    d_tf = capi.DeviceBuffer(plan.output_bytes(taftp.nbytes))
    fr = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F * wl.CHANNEL_BW_HZ, F)
    wp = capi.WeightsPlan(A, B, F, fr)
    d_d = _dev(capi, wl.synth_delays(B, A, 1260 + B, 2e-6, rate_max=1e-9))
    d_w_sep = capi.DeviceBuffer(F * B * A * 2)
    d_w = capi.DeviceBuffer(F * B * A * 2)
    for epoch in (0.0, 321.5):
        wp.generate(d_d, epoch, d_w_sep)
        d_ftpa.fill(0x5A)
        d_w.fill(0x33)
        capi.split_transpose_ib_weights(d_in, d_ftpa, F, plan, d_tf, wp, d_d, epoch, d_w)
        np.testing.assert_array_equal(d_w.download(), d_w_sep.download(), err_msg=f"epoch {epoch}")
        np.testing.assert_array_equal(d_ftpa.download(), oracle.split_transpose(taftp, nhg, A, A, 0, F))
        np.testing.assert_array_equal(d_tf.download(), oracle.ib_tf(oracle.ib_power(taftp, F, ts, 1, nhg, A),
                                                                     off, sc))
    # a weights plan over other channels is refused
    wp2 = capi.WeightsPlan(A, B, 2 * F, np.concatenate([fr, fr]))
    with pytest.raises(capi.FbfError) as ei:
        capi.split_transpose_ib_weights(d_in, d_ftpa, F, plan, d_tf, wp2, d_d, 0.0,
                                        capi.DeviceBuffer(2 * F * B * A * 2))
    assert ei.value.code == capi.FBF_ERR_ARG


def test_split_transpose_ib_rejects_fscrunch(gpu):
    capi = gpu
    A, F, nhg = 8, 8, 1
    plan = capi.IbPlan(A, F, 16, 2, np.zeros(F // 2, np.float32), np.ones(F // 2, np.float32))
    d_in = capi.DeviceBuffer(nhg * A * F * 256 * 4)
    d_ftpa = capi.DeviceBuffer(nhg * A * F * 256 * 4)
    d_tf = capi.DeviceBuffer(plan.output_bytes(nhg * A * F * 256 * 4))
    with pytest.raises(capi.FbfError) as ei:
        capi.split_transpose_ib(d_in, d_ftpa, F, plan, d_tf)
    assert ei.value.code == capi.FBF_ERR_UNSUPPORTED


def test_ib_matches_reference_f32_within_tolerance(gpu, oracle):
    """IncoherentBeamformerTester.cu:124-145 geometry: 64 antennas, 64 channels,
    ts 16, fs 2, 32 heap groups, N(0, 32) voltages, levels 32/32, truncation."""
    capi = gpu
    A, F, nhg, ts, fs = 64, 64, 32, 16, 2
    taftp = gaussian_int8(1100, nhg * A * F * 256 * 4, 32.0, clip=False)
    _, _, ibo, ibs = oracle.power_levels(32.0, 32.0, 44, 16, 4, A, ts, fs)
    Fo = F // fs
    plan = capi.IbPlan(A, F, ts, fs, np.full(Fo, ibo, np.float32), np.full(Fo, ibs, np.float32), quant_mode=1)
    d_in = _dev(capi, taftp)
    d_o = capi.DeviceBuffer(plan.output_bytes(taftp.nbytes))
    plan.beamform(d_in, d_o)
    ref = oracle.ib_reference_f32(taftp, F, ts, fs, nhg, A, ibs, ibo)
    assert np.abs(d_o.download().astype(int) - ref.astype(int)).max() <= 1


# ---------------------------------------------------------------------------
# weights
# ---------------------------------------------------------------------------
def _weights_gpu(capi, d, freqs, epoch, A, B, F):
    plan = capi.WeightsPlan(A, B, F, freqs)
    d_d = _dev(capi, d)
    d_w = capi.DeviceBuffer(F * B * A * 2)
    d_ph = capi.DeviceBuffer(F * B * A * 4)
    plan.generate(d_d, epoch, d_w, d_ph)
    return d_w.download(), d_ph.download(np.float32)


def test_weights_zero_value(gpu, oracle):
    """WeightsManagerTester.cu:95-106: zero delays at epoch 0 -> every weight (127, 0)."""
    A, B, F = 44, 1024, 64
    fr = oracle.channel_frequencies(1.4e9, 56e6, F).astype(np.float64)
    w, _ = _weights_gpu(gpu, np.zeros(A * B * 2, np.float32), fr, 0.0, A, B, F)
    assert np.all(w[0::2] == 127) and np.all(w[1::2] == 0)


def test_weights_real_value(gpu, oracle):
    """WeightsManagerTester.cu:108-119 inputs {rate 1e-11, offset 1e-10}, epoch 10."""
    A, B, F = 44, 1024, 64
    d = np.tile(np.array([1e-11, 1e-10], np.float32), A * B)
    fr = oracle.channel_frequencies(1.4e9, 56e6, F).astype(np.float64)
    w, ph = _weights_gpu(gpu, d, fr, 10.0, A, B, F)
    ref, frac = oracle.weights_f64(d, fr, 10.0, A, B, F, want_phase=True)
    np.testing.assert_array_equal(ph, frac.astype(np.float32))
    diff = np.abs(w.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


@pytest.mark.parametrize("A,B,F", [(5, 7, 13), (3, 100, 9), (64, 5, 1)])
def test_weights_ragged_geometry(gpu, oracle, A, B, F):
    """(beam, antenna) counts that leave a partial work-group and channel counts
    that leave a partial run of channels per thread: every component written,
    the reduced phase equal to the FP64 oracle's rounded to FP32, weights
    within 1 LSB."""
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    d = wl.synth_delays(B, A, 77 + A + B + F, 2e-6, rate_max=1e-9)
    fr = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F * wl.CHANNEL_BW_HZ, F)
    w, ph = _weights_gpu(capi, d, fr, 123.4, A, B, F)
    ref, frac = oracle.weights_f64(d, fr, 123.4, A, B, F, want_phase=True)
    np.testing.assert_array_equal(ph, frac.astype(np.float32))
    assert np.abs(w.astype(int) - ref.astype(int)).max() <= 1


@pytest.mark.parametrize("A,B,F,ntsteps", [(64, 96, 16, 5), (5, 7, 13, 3), (8, 256, 64, 1)])
def test_weights_time_steps(gpu, oracle, A, B, F, ntsteps):
    """generate_weights_k's tstart/tstep/ntsteps (WeightsManager.cu:14-24,
    58-73): step k of the [t][F][B][A] output equals the single-epoch
    generator at tstart + k * tstep, and is within 1 LSB of the FP64 oracle."""
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    d = wl.synth_delays(B, A, 2024 + A + B, 2e-6, rate_max=1e-9)
    fr = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F * wl.CHANNEL_BW_HZ, F)
    plan = capi.WeightsPlan(A, B, F, fr)
    d_d = _dev(capi, d)
    tstart, tstep = 12.25, 0.5360
    d_all = capi.DeviceBuffer(ntsteps * F * B * A * 2)
    d_all.fill(0x77)
    plan.generate_steps(d_d, tstart, tstep, ntsteps, d_all)
    allw = d_all.download().reshape(ntsteps, -1)
    d_one = capi.DeviceBuffer(F * B * A * 2)
    for k in range(ntsteps):
        epoch = tstart + float(k) * tstep
        plan.generate(d_d, epoch, d_one)
        np.testing.assert_array_equal(allw[k], d_one.download(), err_msg=f"step {k}")
        ref = oracle.weights_f64(d, fr, epoch, A, B, F)
        assert np.abs(allw[k].astype(int) - ref.astype(int)).max() <= 1
    with pytest.raises(capi.FbfError):
        plan.generate_steps(d_d, tstart, tstep, ntsteps + 1, d_all)


@pytest.mark.parametrize("epoch", [0.0, 75.0, 300.0, 599.9])
def test_weights_fp64_phase_config4(gpu, oracle, epoch):
    """64 antennas x 1024 beams x 64 channels, tau0 ~ U(-2us, 2us), rate ~ U(-1e-9, 1e-9),
    epochs over 600 s.  Phase argument: the GPU's FP32 reduced phase equals the
    FP64 oracle's rounded to FP32 (max error 2^-25 cycles = 1.9e-7 rad);
    int8 weights within 1 LSB of the FP64 oracle on < 1e-3 of the components."""
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    A, B, F = 64, 1024, 64
    d = wl.synth_delays(B, A, 4242, 2e-6, rate_max=1e-9)
    fr = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F * wl.CHANNEL_BW_HZ, F)
    w, ph = _weights_gpu(capi, d, fr, epoch, A, B, F)
    ref, frac = oracle.weights_f64(d, fr, epoch, A, B, F, want_phase=True)
    err_rad = np.abs(ph.astype(np.float64) - frac).max() * 2 * np.pi
    diff = np.abs(w.astype(int) - ref.astype(int))
    print(f"weights 64x1024x64 epoch {epoch}: max phase error {err_rad:.3e} rad, "
          f"components 1 LSB off the FP64 oracle {(diff > 0).mean():.3e}, max diff {diff.max()}")
    assert err_rad <= 2 * np.pi * 2.0 ** -25
    assert diff.max() <= 1
    assert (diff > 0).mean() < 1e-3


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11])
@pytest.mark.parametrize("F,T,A,B,ts,fs", [(4, 512, 64, 512, 16, 1), (4, 256, 44, 96, 4, 2), (2, 512, 8, 64, 16, 1),
                                            (2, 256, 40, 48, 64, 1), (2, 256, 128, 32, 8, 1),
                                            (4, 512, 40, 96, 4, 4), (4, 1024, 64, 256, 64, 2), (2, 256, 16, 48, 8, 1),
                                            (8, 256, 24, 72, 32, 4), (4, 512, 64, 136, 16, 2),
                                            (4, 1536, 64, 256, 8, 1), (3, 1040, 48, 200, 16, 1)])
def test_cb_every_variant_bit_exact(gpu, oracle, variant, F, T, A, B, ts, fs):
    """Every kernel variant (MFMA shape x tile) gives the oracle's sums and power
    (the transposed variants 8 and 9 serve int8 and power output and fall back
    to the LDS variant for sums and for geometries they do not cover)."""
    capi = gpu
    ftpa = gaussian_int8(1300 + A + variant, F * T * 2 * A * 2, 16.0)
    w = unit_weights(1400 + B, F * B * A)
    off, sc = _levels(oracle, A, ts, fs, F)
    plan = capi.CbPlan(A, B, F, ts, fs, off, sc, nsamples_per_heap_out=T // ts)
    plan.set_variant(variant)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_s = capi.DeviceBuffer(F * T * 2 * B * 2 * 4)
    plan.beam_sums(d_v, d_w, d_s)
    np.testing.assert_array_equal(d_s.download(np.int32), oracle.cb_sums(ftpa, w, F, T, A, B).reshape(-1))
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, fs)
    d_o = capi.DeviceBuffer((F // fs) * (T // ts) * B)
    plan.beamform(d_v, d_w, d_o)
    np.testing.assert_array_equal(d_o.download(), oracle.cb_tbtf(p, off, sc, T // ts))
    d_p = capi.DeviceBuffer((F // fs) * (T // ts) * B * 8)
    plan.beam_power(d_v, d_w, d_p)
    np.testing.assert_array_equal(d_p.download(np.float64).reshape(F // fs, T // ts, B), p)


@pytest.mark.parametrize("variant", [5, 8, 9, 10, 11])
@pytest.mark.parametrize("mode", [0, 1])
def test_cb_quantiser_rounding_boundaries(gpu, oracle, variant, mode):
    """Levels that put many outputs exactly on (or next to) rounding
    boundaries: small integer voltages with unit weights make the power small
    integers, and offset/scale of 0/2, 1/4, 0.5/1, 0/1 give exact
    half-integers and integers; a zero and a negative scale cover the
    non-finite and sign cases.  Every variant must match the oracle's
    IEEE-division quantiser exactly."""
    capi = gpu
    F, T, A, B, ts = 8, 256, 8, 32, 4
    rng = np.random.default_rng(77 + variant)
    ftpa = rng.integers(-2, 3, F * T * 2 * A * 2).astype(np.int8)
    w = np.zeros(F * B * A * 2, np.int8)
    w[0::2] = 1
    off = np.array([0.0, 1.0, 0.5, 0.0, 3.0, 0.0, 40.0, 0.25], np.float32)
    sc = np.array([2.0, 4.0, 1.0, 1.0, 0.0, -2.0, 0.5, 1e-40], np.float32)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts, quant_mode=mode)
    plan.set_variant(variant)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_o = capi.DeviceBuffer(F * (T // ts) * B)
    plan.beamform(d_v, d_w, d_o)
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, 1)
    ref = oracle.cb_tbtf(p, off, sc, T // ts, mode)
    np.testing.assert_array_equal(d_o.download(), ref)
    # the levels really produce ties: half-integer quotients in channel 0
    assert np.any(p[0] % 2 == 1)


def test_cb_f32_estimate_at_its_bound(gpu, oracle):
    """Levels with |offset / scale| just below 2^20, the bound under which the
    transposed kernels estimate the quantiser's quotient in FP32 (two channels
    just above it take the FP64 estimate).  Each channel's offset is one of
    its own powers (that output's quotient is 0) and its scale is offset / r,
    so only powers within 128 * scale of the offset (about 1e-4 relative)
    are not saturated: a few per channel, the rest exercise saturation at
    both ends.  Both rounding modes, the kernel with 64 and with 128 beams
    per wave; every output must match the oracle."""
    capi = gpu
    F, T, A, B, ts = 16, 4096, 32, 64, 16
    rng = np.random.default_rng(2020)
    ftpa = gaussian_int8(2021, F * T * 2 * A * 2, 24.0)
    w = rng.integers(-127, 128, F * B * A * 2).astype(np.int8)
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, 1)
    r = np.concatenate([np.linspace(0.999e6, 1.0485e6, F - 2), [1.0487e6, 1.2e6]])
    off = np.array([np.median(p[f]) for f in range(F)], np.float32)
    off = np.array([p[f].reshape(-1)[np.argmin(np.abs(p[f].reshape(-1) - off[f]))] for f in range(F)], np.float32)
    sc = (off.astype(np.float64) / r).astype(np.float32)
    below = np.abs(off / sc) < np.float32(2 ** 20)
    assert below[:F - 2].all() and not below[F - 2:].any()
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_o = capi.DeviceBuffer(F * (T // ts) * B)
    for mode in (0, 1):
        ref = oracle.cb_tbtf(p, off, sc, T // ts, mode)
        inside = np.count_nonzero((ref > -128) & (ref < 127))
        assert inside >= 2 * F   # the unsaturated band is reached
        for variant in (0, 9, 11):
            plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts, quant_mode=mode)
            plan.set_variant(variant)
            plan.beamform(d_v, d_w, d_o)
            np.testing.assert_array_equal(d_o.download(), ref, err_msg=f"mode {mode} variant {variant}")


@pytest.mark.parametrize("F,T,A,B,ts,mode", [
    (4, 2048, 64, 512, 16, 0),     # 8 units: time ranges, one unit per work-group
    (37, 256, 16, 576, 16, 1),     # 333 units: two per work-group, the last work-group holds one
    (64, 128, 8, 1536, 8, 0),      # 1536 units: six per work-group
    (2, 4096, 56, 64, 4, 1),       # K = 112 (pieces past K), window 8: one tile per pass
    (3, 1200, 40, 72, 8, 0),       # KSTEPS 3, a partial last chunk, a partial beam group
    (5, 512, 24, 200, 64, 1),      # window 128: four tiles per window, two windows per pass
    (2, 8192, 8, 64, 32, 0),       # KSTEPS 1: eight-tile rings
])
def test_cb_wave_private_staging_bit_exact(gpu, oracle, F, T, A, B, ts, mode):
    """Variant 10 (wave-private staging, dynamic chunks): unit counts that put
    one, two (with a ragged last work-group) and six units on a work-group,
    time ranges when units are few, every K step count, window sizes from one
    tile per pass to several tiles per window, and ragged tiles, chunks and
    beam groups; int8 TBTF bit-exact with the oracle, also read in place
    from a wider FTPA block (subarray view)."""
    capi = gpu
    rng = np.random.default_rng(5000 + F * 7 + A)
    ftpa = gaussian_int8(5000 + F * 7 + A, F * T * 2 * A * 2, 24.0)
    w = rng.integers(-127, 128, F * B * A * 2).astype(np.int8)
    off, sc = _levels(oracle, A, ts, 1, F)
    ohs = T // ts
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=ohs, quant_mode=mode)
    plan.set_variant(10)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_o = capi.DeviceBuffer(F * (T // ts) * B)
    plan.beamform(d_v, d_w, d_o)
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, 1)
    ref = oracle.cb_tbtf(p, off, sc, ohs, mode)
    np.testing.assert_array_equal(d_o.download(), ref)
    # the same antennas as a slice of a wider block
    A2, a0 = A + 16, 8
    blk = rng.integers(-127, 128, F * T * 2 * A2 * 2).astype(np.int8).reshape(F, T * 2, A2 * 2)
    blk[:, :, 2 * a0:2 * (a0 + A)] = ftpa.reshape(F, T * 2, A * 2)
    d_b = _dev(capi, blk.reshape(-1))
    d_o2 = capi.DeviceBuffer(F * (T // ts) * B)
    plan.beamform_subarray(d_b, A2, a0, d_w, d_o2)
    np.testing.assert_array_equal(d_o2.download(), ref)


@pytest.mark.parametrize("variant", [0, 5, 9, 10, 11])
def test_cb_set_levels_applies_to_later_blocks(gpu, oracle, variant):
    """New levels on a live plan (PipelineConfig's output level changing
    between blocks) apply to the next beamform call: the quantiser's
    thresholds are rebuilt with them, including after a switch to levels the
    thresholds do not cover (zero scale) and back."""
    capi = gpu
    F, T, A, B, ts = 8, 512, 16, 64, 16
    rng = np.random.default_rng(91 + variant)
    ftpa = rng.integers(-40, 41, F * T * 2 * A * 2).astype(np.int8)
    w = rng.integers(-100, 101, F * B * A * 2).astype(np.int8)
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, 1)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    d_o = capi.DeviceBuffer(F * (T // ts) * B)
    off1, sc1 = _levels(oracle, A, ts, 1, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off1, sc1, nsamples_per_heap_out=T // ts)
    plan.set_variant(variant)
    # the last set has |offset / scale| above 2^20 in half the channels,
    # where the beamformer's quantiser takes its FP64 estimate
    levels = [(off1, sc1),
              (off1 * 1.5, sc1 * 0.25),
              (off1, np.where(np.arange(F) % 3 == 0, 0.0, sc1).astype(np.float32)),
              (off1 * 0.5, sc1 * 3.0),
              (off1, np.where(np.arange(F) % 2 == 0, sc1 * 1e-5, sc1).astype(np.float32))]
    for k, (off, sc) in enumerate(levels):
        if k:
            plan.set_levels(off, sc)
        plan.beamform(d_v, d_w, d_o)
        np.testing.assert_array_equal(d_o.download(), oracle.cb_tbtf(p, off, sc, T // ts), err_msg=f"levels {k}")


def test_cb_set_levels_waits_for_work_in_flight(gpu, oracle):
    """New levels set while beamforms enqueued on a library stream may still
    run: those calls finish with the old levels (the update waits for the
    device), and a call enqueued after it uses the new ones."""
    capi = gpu
    F, T, A, B, ts = 16, 8192, 64, 512, 16
    ftpa = gaussian_int8(95, F * T * 2 * A * 2, 16.0)
    w = unit_weights(96, F * B * A)
    off1, sc1 = _levels(oracle, A, ts, 1, F)
    off2, sc2 = off1 * 1.25, sc1 * 0.5
    plan = capi.CbPlan(A, B, F, ts, 1, off1, sc1)
    d_v, d_w = _dev(capi, ftpa), _dev(capi, w)
    nout = plan.output_bytes(ftpa.nbytes)
    outs = [capi.DeviceBuffer(nout) for _ in range(4)]
    s = capi.Stream()
    for o in outs[:3]:
        plan.beamform(d_v, d_w, o, stream=s)
    plan.set_levels(off2, sc2)
    plan.beamform(d_v, d_w, outs[3], stream=s)
    s.sync()
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, 1)
    ohs = 8192 // F
    ref1 = oracle.cb_tbtf(p, off1, sc1, ohs)
    for k in range(3):
        np.testing.assert_array_equal(outs[k].download(), ref1, err_msg=f"call {k} (old levels)")
    np.testing.assert_array_equal(outs[3].download(), oracle.cb_tbtf(p, off2, sc2, ohs))
