"""Batched uneven subarrays (one block split 40/16/8 antennas with 768/256/64
beams), the subarray entry point and the autotuner, against the oracle."""
import numpy as np
import pytest

from conftest import gaussian_int8, unit_weights

pytestmark = pytest.mark.gpu

A_TOT = 64
SPLIT = [dict(start_antenna=0, nantennas=40, nbeams=768), dict(start_antenna=40, nantennas=16, nbeams=256),
         dict(start_antenna=56, nantennas=8, nbeams=64)]


def _levels(oracle, A, ts, fs, F):
    cbo, cbs, _, _ = oracle.power_levels(16.0, 24.0, A, ts, fs, A_TOT, 16, 1)
    return np.full(F // fs, cbo, np.float32), np.full(F // fs, cbs, np.float32)


def _reference(oracle, taftp, nhg, F, sd, ts, fs, off, sc, ohs, w):
    A, a0, B = sd["nantennas"], sd["start_antenna"], sd["nbeams"]
    ftpa = oracle.split_transpose(taftp, nhg, A_TOT, A, a0, F)
    T = nhg * 256
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, fs)
    return oracle.cb_tbtf(p, off, sc, ohs)


@pytest.mark.parametrize("ts", [4, 16, 64])
def test_cb_subarray_view_equals_split_then_beamform(gpu, oracle, ts):
    capi = gpu
    F, nhg = 8, 2
    T = nhg * 256
    taftp = gaussian_int8(3000 + ts, nhg * A_TOT * F * 256 * 4, 16.0)
    d_taftp = capi.DeviceBuffer.from_array(taftp)
    d_full = capi.DeviceBuffer(taftp.nbytes)
    capi.split_transpose(d_taftp, d_full, A_TOT, A_TOT, 0, F)
    for k, sd in enumerate(SPLIT):
        A, B = sd["nantennas"], sd["nbeams"]
        w = unit_weights(3100 + k, F * B * A)
        off, sc = _levels(oracle, A, ts, 1, F)
        plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts)
        d_w = capi.DeviceBuffer.from_array(w)
        d_o = capi.DeviceBuffer((T // ts) * B * F)
        plan.beamform_subarray(d_full, A_TOT, sd["start_antenna"], d_w, d_o)
        ref = _reference(oracle, taftp, nhg, F, sd, ts, 1, off, sc, T // ts, w)
        np.testing.assert_array_equal(d_o.download(), ref)


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("serial", [False, True])
@pytest.mark.parametrize("ib_fs", [1, 2])
def test_batch_40_16_8_matches_oracle(gpu, oracle, use_graph, serial, ib_fs):
    """Every subarray's output (and the incoherent beam) equals the oracle, with
    the subarrays side by side or in turn on one stream, and with the
    incoherent beam fused into the transpose (fscrunch 1) or on its own
    (fscrunch 2).  With the graph, it is replayed several times with eager
    work and host syncs in between, and every replay is checked."""
    capi = gpu
    F, nhg, ts = 8, 2, 16
    T = nhg * 256
    subs = [dict(sd, tscrunch=ts, fscrunch=1, nsamples_per_heap_out=T // ts) for sd in SPLIT]
    levels = [_levels(oracle, sd["nantennas"], ts, 1, F) for sd in subs]
    batch = capi.Batch(A_TOT, F, nhg, subs, levels)
    batch.set_serial(serial)
    assert batch.serial == serial
    _, _, ibo, ibs = oracle.power_levels(16.0, 24.0, 8, 16, 1, A_TOT, 16, ib_fs)
    Fi = F // ib_fs
    batch.enable_ib(16, ib_fs, np.full(Fi, ibo, np.float32), np.full(Fi, ibs, np.float32))
    ws = [unit_weights(3200 + k, F * sd["nbeams"] * sd["nantennas"]) for k, sd in enumerate(subs)]
    d_ws = [capi.DeviceBuffer.from_array(w) for w in ws]
    outs = [capi.DeviceBuffer(batch.output_bytes(i)) for i in range(len(subs) + 1)]
    s = capi.Stream()
    scratch = capi.DeviceBuffer(1 << 20)
    for rep in range(3):
        taftp = gaussian_int8(3300 + rep, nhg * A_TOT * F * 256 * 4, 16.0)
        d_taftp = capi.DeviceBuffer.from_array(taftp) if rep == 0 else d_taftp
        if rep:
            d_taftp.upload(taftp, s)
        for o in outs:
            o.fill(0x55, s)
        batch.run(d_taftp, d_ws, outs, use_graph=use_graph, stream=s)
        s.sync()
        scratch.fill(rep, s)          # eager work between replays
        s.sync()
        for k, sd in enumerate(subs):
            ref = _reference(oracle, taftp, nhg, F, sd, ts, 1, levels[k][0], levels[k][1], T // ts, ws[k])
            np.testing.assert_array_equal(outs[k].download(), ref, err_msg=f"replay {rep} subarray {k}")
        pib = oracle.ib_power(taftp, F, 16, ib_fs, nhg, A_TOT)
        np.testing.assert_array_equal(outs[-1].download(),
                                      oracle.ib_tf(pib, np.full(Fi, ibo, np.float32), np.full(Fi, ibs, np.float32)))


def test_autotune_picks_a_variant_that_is_bit_exact(gpu, oracle):
    capi = gpu
    F, T, A, B, ts = 4, 512, 40, 768, 16
    off, sc = _levels(oracle, A, ts, 1, F)
    plan = capi.CbPlan(A, B, F, ts, 1, off, sc, nsamples_per_heap_out=T // ts)
    var, chunk, ms = plan.autotune(T, reps=2)
    print("autotune:", var, chunk, ms)
    assert var >= 1 and ms > 0
    ftpa = gaussian_int8(3400, F * T * 2 * A * 2, 16.0)
    w = unit_weights(3401, F * B * A)
    d_v, d_w = capi.DeviceBuffer.from_array(ftpa), capi.DeviceBuffer.from_array(w)
    d_o = capi.DeviceBuffer((T // ts) * B * F)
    plan.beamform(d_v, d_w, d_o)
    p = oracle.cb_power(ftpa, w, F, T, A, B, ts, 1)
    np.testing.assert_array_equal(d_o.download(), oracle.cb_tbtf(p, off, sc, T // ts))
    # a new plan of the same geometry starts from the tuned choice, and chunk
    # lengths below the default are exact too
    for c in (1, 2, 4):
        plan.set_variant(5)
        plan.set_chunk(c)
        d_o.fill(0)
        plan.beamform(d_v, d_w, d_o)
        np.testing.assert_array_equal(d_o.download(), oracle.cb_tbtf(p, off, sc, T // ts))
