"""Serving engine: continuous batching of blocks from uneven feeds.

Each result is the same alone and in the mix (bit-exact, and equal to the
oracle), admission respects the HBM budget, weights refreshed mid-stream
apply to the blocks admitted after the refresh, and device-resident input
works."""
import numpy as np
import pytest

from conftest import gaussian_int8

pytestmark = pytest.mark.gpu

# (antennas, beams, channels) of four uneven feeds, 4 heap groups per block
FEEDS = [(64, 128, 16), (32, 64, 16), (16, 32, 8), (8, 16, 8)]
NHG = 4
TS = 16


def _feed_levels(oracle, A, F, ts=TS, fs=1):
    cbo, cbs, ibo, ibs = oracle.power_levels(16.0, 24.0, A, ts, fs, A, 16, 1)
    return ((np.full(F // fs, cbo, np.float32), np.full(F // fs, cbs, np.float32)),
            (np.full(F, ibo, np.float32), np.full(F, ibs, np.float32)))


def _expected(oracle, taftp, A, B, F, w, cbl, ibl):
    T = NHG * 256
    ftpa = oracle.split_transpose(taftp, NHG, A, A, 0, F)
    p = oracle.cb_power(ftpa, w, F, T, A, B, TS, 1)
    cb = oracle.cb_tbtf(p, cbl[0], cbl[1], T // TS)
    pib = oracle.ib_power(taftp, F, 16, 1, NHG, A)
    return cb, oracle.ib_tf(pib, ibl[0], ibl[1])


def _make_server(capi, oracle, budget=1 << 30, nslots=4):
    srv = capi.Server(nslots, budget)
    info = []
    for k, (A, B, F) in enumerate(FEEDS):
        cbl, ibl = _feed_levels(oracle, A, F)
        fid = srv.add_feed(A, B, F, NHG, TS, 1, cbl, ib=(16, 1), ib_levels=ibl, nsamples_per_heap_out=NHG * 256 // TS)
        delays = np.zeros((B * A, 2), np.float32)
        delays[:, 1] = capi.synth_uniform_f32(600 + k, -1e-6, 1e-6, B * A)
        freqs = capi.channel_frequencies_f64(1284e6, F * 208984.375, F)
        srv.set_delays(fid, delays.reshape(-1), freqs, 0.0)
        w = oracle.weights_f64(delays.reshape(-1), freqs, 0.0, A, B, F)
        info.append(dict(fid=fid, A=A, B=B, F=F, cbl=cbl, ibl=ibl, delays=delays, freqs=freqs, w_oracle=w))
    srv.start()
    return srv, info


def _weights_on_device(capi, info_k):
    """The weights the server generated, recovered through a plain weights plan
    (the same kernel and inputs)."""
    A, B, F = info_k["A"], info_k["B"], info_k["F"]
    wp = capi.WeightsPlan(A, B, F, info_k["freqs"])
    d_d = capi.DeviceBuffer.from_array(info_k["delays"].reshape(-1))
    d_w = capi.DeviceBuffer(F * B * A * 2)
    wp.generate(d_d, 0.0, d_w)
    return d_w.download()


def _collect(srv, n):
    got = {}
    for _ in range(n):
        r = srv.poll(timeout_ms=20000)
        assert r is not None, "block did not complete"
        tag, status, lat = r
        assert status == 0
        got[tag] = lat
    return got


def test_each_result_same_alone_and_in_the_mix(gpu, oracle):
    capi = gpu
    srv, info = _make_server(capi, oracle)
    inputs, outs = {}, {}
    tag = 0
    # alone: one block of feed 0
    k0 = info[0]
    bb, cbb, ibb = srv.feed_bytes(k0["fid"])
    x = capi.HostBuffer(bb)
    x.array[:] = gaussian_int8(4000, bb, 16.0)
    alone_cb, alone_ib = capi.HostBuffer(cbb), capi.HostBuffer(ibb)
    srv.submit(k0["fid"], x, alone_cb, alone_ib, tag=999)
    _collect(srv, 1)
    a_cb, a_ib = alone_cb.array.copy(), alone_ib.array.copy()
    # in the mix: 6 blocks of every feed submitted back to back, the same
    # feed-0 block among them
    for rep in range(6):
        for k, fi in enumerate(info):
            bb, cbb, ibb = srv.feed_bytes(fi["fid"])
            if k == 0 and rep == 3:
                inp = x
            else:
                inp = capi.HostBuffer(bb)
                inp.array[:] = gaussian_int8(4100 + 10 * rep + k, bb, 16.0)
            cb_o, ib_o = capi.HostBuffer(cbb), capi.HostBuffer(ibb)
            srv.submit(fi["fid"], inp, cb_o, ib_o, tag=tag)
            inputs[tag] = (k, inp)
            outs[tag] = (cb_o, ib_o)
            tag += 1
    _collect(srv, len(inputs))
    mix_tag = [t for t, (k, inp) in inputs.items() if inp is x][0]
    np.testing.assert_array_equal(outs[mix_tag][0].array, a_cb)
    np.testing.assert_array_equal(outs[mix_tag][1].array, a_ib)
    # and every block equals the oracle
    wdev = [_weights_on_device(capi, fi) for fi in info]
    for t, (k, inp) in inputs.items():
        fi = info[k]
        assert np.abs(wdev[k].astype(int) - fi["w_oracle"].astype(int)).max() <= 1
        cb, ib = _expected(oracle, inp.array, fi["A"], fi["B"], fi["F"], wdev[k], fi["cbl"], fi["ibl"])
        np.testing.assert_array_equal(outs[t][0].array, cb, err_msg=f"tag {t}")
        np.testing.assert_array_equal(outs[t][1].array, ib, err_msg=f"tag {t}")
    st = srv.stats()
    assert st["completed"] == 25 and st["failed"] == 0 and st["p99_ms"] > 0
    srv.close()


def test_admission_respects_hbm_budget(gpu, oracle):
    """A budget that holds one block of the largest feed at a time: every block
    still completes correctly and the arena's high-water mark stays in budget."""
    capi = gpu
    A, B, F = FEEDS[0]
    block = NHG * A * F * 256 * 4
    budget = 2 * block + (1 << 20) + 128 * 1024 + 2 * 256 + 64 * 1024   # input + FTPA + staging + outputs of one block
    srv = capi.Server(8, budget)
    cbl, ibl = _feed_levels(oracle, A, F)
    fid = srv.add_feed(A, B, F, NHG, TS, 1, cbl, ib=(16, 1), ib_levels=ibl, nsamples_per_heap_out=(NHG * 256) // TS)
    w = gaussian_int8(4200, F * B * A * 2, 60.0)
    srv.set_weights(fid, w)
    srv.start()
    bb, cbb, ibb = srv.feed_bytes(fid)
    ins, outs = [], []
    for i in range(6):
        h = capi.HostBuffer(bb)
        h.array[:] = gaussian_int8(4300 + i, bb, 16.0)
        o = (capi.HostBuffer(cbb), capi.HostBuffer(ibb))
        srv.submit(fid, h, o[0], o[1], tag=i)
        ins.append(h)
        outs.append(o)
    _collect(srv, 6)
    st = srv.stats()
    assert st["hbm_peak"] <= budget
    for i in range(6):
        cb, ib = _expected(oracle, ins[i].array, A, B, F, w, cbl, ibl)
        np.testing.assert_array_equal(outs[i][0].array, cb)
    srv.close()


def test_weights_refresh_applies_to_later_blocks_and_device_input(gpu, oracle):
    capi = gpu
    A, B, F = FEEDS[2]
    srv = capi.Server(2, 1 << 28)
    cbl, ibl = _feed_levels(oracle, A, F)
    T = NHG * 256
    fid = srv.add_feed(A, B, F, NHG, TS, 1, cbl, ib=(16, 1), ib_levels=ibl, nsamples_per_heap_out=T // TS)
    w1 = gaussian_int8(4400, F * B * A * 2, 60.0)
    w2 = gaussian_int8(4401, F * B * A * 2, 60.0)
    srv.set_weights(fid, w1)
    srv.start()
    bb, cbb, ibb = srv.feed_bytes(fid)
    x = gaussian_int8(4402, bb, 16.0)
    d_x = capi.DeviceBuffer.from_array(x)
    o1 = (capi.HostBuffer(cbb), capi.HostBuffer(ibb))
    srv.submit(fid, d_x, o1[0], o1[1], tag=1, on_device=True)
    _collect(srv, 1)
    srv.set_weights(fid, w2)
    o2 = (capi.HostBuffer(cbb), capi.HostBuffer(ibb))
    srv.submit(fid, d_x, o2[0], o2[1], tag=2, on_device=True)
    _collect(srv, 1)
    cb1, _ = _expected(oracle, x, A, B, F, w1, cbl, ibl)
    cb2, _ = _expected(oracle, x, A, B, F, w2, cbl, ibl)
    np.testing.assert_array_equal(o1[0].array, cb1)
    np.testing.assert_array_equal(o2[0].array, cb2)
    assert np.mean(cb1 != cb2) > 0.5
    srv.close()


def test_poisson_driver_small_feeds(gpu):
    """The load generator against the engine: four feeds of 8 heap groups,
    Poisson arrivals at 1.5x real time for one second; every block completes."""
    from psrdada_cpp_amd import serving
    feeds = serving.four_subarray_feeds(8)
    srv, fids = serving.make_server(feeds, 2 << 30, nslots=4)
    try:
        r = serving.run_poisson(srv, fids, feeds, 1.5, 1.0)
        print(r)
        assert r["failed"] == 0 and r["blocks"] > 100
        assert r["hbm_peak"] <= r["hbm_budget"]
        s = serving.run_saturation(srv, fids, feeds, 8, device_input=True)
        print(s)
        assert s["blocks"] == 32 and s["x_real_time"] > 1.0
    finally:
        srv.close()


@pytest.mark.parametrize("nheap_groups,seconds", [(32, 4.0), (8, 2.0)])
def test_poisson_p99_within_one_block(gpu, nheap_groups, seconds):
    """The serving target: four feeds (64/768/64, 32/256/64, 16/128/32,
    8/64/16 antennas/beams/channels), Poisson arrivals at 1.5x real time per
    feed from pinned host memory (PCIe-inclusive), after one untimed block per
    feed: every block completes and the p99 block latency, submit to the
    outputs back in host memory, stays below one block duration (39.2 ms at
    32 heap groups, the config-5 block; 9.8 ms at 8)."""
    from psrdada_cpp_amd import serving
    feeds = serving.four_subarray_feeds(nheap_groups)
    block_ms = feeds[0].block_seconds * 1e3
    srv, fids = serving.make_server(feeds, 16 << 30, nslots=8)
    try:
        r = serving.run_poisson(srv, fids, feeds, 1.5, seconds)
    finally:
        srv.close()
    print(f"{nheap_groups} heap groups ({block_ms:.1f} ms blocks):", r)
    assert r["failed"] == 0 and r["blocks"] >= 4 * seconds * 1.5 / feeds[0].block_seconds * 0.7
    assert r["latency_ms"]["p99"] < block_ms, r["latency_ms_per_feed"]
