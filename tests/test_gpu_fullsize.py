"""Full-size parity: the exact launches the benchmark times, bit for bit
against the oracle (CoherentBeamformerTester.cu:148-185 compares whole
blocks the same way).

* Config 2: one 64-antenna, 64-channel, 32-heap-group block (8192 spectra),
  512 beams, tscrunch 16, with the benchmark's seed.  The benchmark's
  default step (test_config2_one_pass_blocks_bit_exact): both beams in one
  pass over TAFTP, the reorder's launch generating the next block's
  weights (fbf_cb_ib_beamform_taftp_next), over two blocks, the second
  beamformed with the weights the first left.  The split-transpose path
  (test_config2_block_bit_exact): the split-transpose + incoherent beam
  launch that also generates the block's weights
  (fbf_split_transpose_ib_weights), then the coherent beamformer.  The
  coherent kernel's one-round grid partition depends on F, T and the
  compute-unit count only, so these are the launches the benchmark times.
* Config 3: the same block split 40/16/8 antennas with 768/256/64 beams plus
  the incoherent beam, as one captured graph after the autotuner, for
  tscrunch 4, 16 and 64; replayed on two blocks with eager work in between,
  every replay checked.
* Config 5: one block of each of the four serving feeds (64/768/64,
  32/256/64, 16/128/32, 8/64/16 antennas/beams/channels) at 32 heap groups
  through the serving engine.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

BENCH_SEED = 20260101


def _weights(capi, wl, cfg, sub, seed, epoch=0.0):
    """Delay-generated weights of one subarray (the benchmark's inputs), on
    the device and as bytes."""
    A, B, F = sub.nantennas, sub.nbeams, cfg.nchans
    d_d = capi.DeviceBuffer.from_array(wl.synth_delays(B, A, seed, cfg.tau_max_s))
    wp = capi.WeightsPlan(A, B, F, cfg.channel_frequencies())
    d_w = capi.DeviceBuffer(cfg.weights_bytes(sub))
    wp.generate(d_d, epoch, d_w)
    return d_w, d_w.download()


@pytest.mark.parametrize("quant_mode", [0, 1])
def test_config2_block_bit_exact(gpu, oracle, quant_mode):
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    cfg = wl.config2()
    sub = cfg.subarrays[0]
    A, F, B, T, nhg = sub.nantennas, cfg.nchans, sub.nbeams, cfg.nsamples, cfg.nheap_groups
    print(f"config 2 block: {A} ant, {F} ch, {T} spectra, {B} beams, ts {sub.tscrunch}, quant {quant_mode}",
          flush=True)
    taftp = wl.synth_taftp(cfg, BENCH_SEED)
    (cb_off, cb_sc), (ib_off, ib_sc) = cfg.levels(sub)
    s = capi.Stream()
    d_taftp = capi.DeviceBuffer.from_array(taftp)
    d_ftpa = capi.DeviceBuffer(cfg.ftpa_bytes(sub))
    delays = wl.synth_delays(B, A, BENCH_SEED + 1, cfg.tau_max_s)
    d_delays = capi.DeviceBuffer.from_array(delays)
    wp = capi.WeightsPlan(A, B, F, cfg.channel_frequencies())
    d_w = capi.DeviceBuffer(cfg.weights_bytes(sub))
    epoch = cfg.block_seconds * 7     # the 8th block of a run
    cb = capi.CbPlan(A, B, F, sub.tscrunch, sub.fscrunch, cb_off, cb_sc, nsamples_per_heap_out=cfg.cb_ohs(sub),
                     quant_mode=quant_mode)
    ib = capi.IbPlan(cfg.total_nantennas, F, cfg.ib_tscrunch, cfg.ib_fscrunch, ib_off, ib_sc, quant_mode=quant_mode)
    d_cb = capi.DeviceBuffer(cfg.cb_out_bytes(sub))
    d_ib = capi.DeviceBuffer(cfg.ib_out_bytes())
    d_ftpa.fill(0x5A, s)
    d_cb.fill(0x5A, s)
    d_w.fill(0x5A, s)
    capi.split_transpose_ib_weights(d_taftp, d_ftpa, F, ib, d_ib, wp, d_delays, epoch, d_w, stream=s)
    cb.beamform(d_ftpa, d_w, d_cb, stream=s)
    s.sync()
    w = d_w.download()

    w_ref = oracle.weights_f64(delays, cfg.channel_frequencies(), epoch, A, B, F)
    dw = np.abs(w.astype(int) - w_ref.astype(int))
    assert dw.max() <= 1 and (dw > 0).mean() < 1e-3
    ftpa = oracle.split_transpose(taftp, nhg, A, A, 0, F)
    np.testing.assert_array_equal(d_ftpa.download(), ftpa)
    pib = oracle.ib_power(taftp, F, cfg.ib_tscrunch, cfg.ib_fscrunch, nhg, A, nthreads=16)
    np.testing.assert_array_equal(d_ib.download(), oracle.ib_tf(pib, ib_off, ib_sc, quant_mode))
    p = oracle.cb_power(ftpa, w, F, T, A, B, sub.tscrunch, sub.fscrunch, nthreads=16)
    got = d_cb.download()
    np.testing.assert_array_equal(got, oracle.cb_tbtf(p, cb_off, cb_sc, cfg.cb_ohs(sub), quant_mode))
    assert got.astype(float).std() > 3.0 and np.mean(np.abs(got.astype(int)) == 127) < 0.01


@pytest.mark.parametrize("quant_mode", [0, 1])
def test_config2_one_pass_blocks_bit_exact(gpu, oracle, quant_mode):
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    cfg = wl.config2()
    sub = cfg.subarrays[0]
    A, F, B, T, nhg = sub.nantennas, cfg.nchans, sub.nbeams, cfg.nsamples, cfg.nheap_groups
    print(f"config 2 one-pass blocks: {A} ant, {F} ch, {T} spectra, {B} beams, ts {sub.tscrunch}, quant {quant_mode}",
          flush=True)
    (cb_off, cb_sc), (ib_off, ib_sc) = cfg.levels(sub)
    s = capi.Stream()
    delays = wl.synth_delays(B, A, BENCH_SEED + 1, cfg.tau_max_s)
    d_delays = capi.DeviceBuffer.from_array(delays)
    wp = capi.WeightsPlan(A, B, F, cfg.channel_frequencies())
    cb = capi.CbPlan(A, B, F, sub.tscrunch, sub.fscrunch, cb_off, cb_sc, nsamples_per_heap_out=cfg.cb_ohs(sub),
                     quant_mode=quant_mode)
    ib = capi.IbPlan(cfg.total_nantennas, F, cfg.ib_tscrunch, cfg.ib_fscrunch, ib_off, ib_sc, quant_mode=quant_mode)
    assert cb.ib_fused(ib, T)
    d_w = capi.DeviceBuffer(cfg.weights_bytes(sub))
    d_cb = capi.DeviceBuffer(cfg.cb_out_bytes(sub))
    d_ib = capi.DeviceBuffer(cfg.ib_out_bytes())
    epoch = cfg.block_seconds * 7
    wp.generate(d_delays, epoch, d_w, stream=s)
    for blk, seed in enumerate((BENCH_SEED, BENCH_SEED + 101)):
        taftp = wl.synth_taftp(cfg, seed)
        d_taftp = capi.DeviceBuffer.from_array(taftp)
        s.sync()
        w = d_w.download()   # this block's weights
        w_ref = oracle.weights_f64(delays, cfg.channel_frequencies(), epoch, A, B, F)
        dw = np.abs(w.astype(int) - w_ref.astype(int))
        assert dw.max() <= 1 and (dw > 0).mean() < 1e-3, f"block {blk} weights"
        d_cb.fill(0x5A, s)
        d_ib.fill(0x5A, s)
        cb.ib_beamform_taftp_next(ib, d_taftp, d_w, d_cb, d_ib, wp, d_delays, epoch + cfg.block_seconds, d_w,
                                  stream=s)
        s.sync()
        ftpa = oracle.split_transpose(taftp, nhg, A, A, 0, F)
        pib = oracle.ib_power(taftp, F, cfg.ib_tscrunch, cfg.ib_fscrunch, nhg, A, nthreads=16)
        np.testing.assert_array_equal(d_ib.download(), oracle.ib_tf(pib, ib_off, ib_sc, quant_mode))
        p = oracle.cb_power(ftpa, w, F, T, A, B, sub.tscrunch, sub.fscrunch, nthreads=16)
        np.testing.assert_array_equal(d_cb.download(), oracle.cb_tbtf(p, cb_off, cb_sc, cfg.cb_ohs(sub), quant_mode))
        d_taftp.free()
        epoch += cfg.block_seconds
    # The weights left for the next block equal the generator's at that epoch.
    d_chk = capi.DeviceBuffer(cfg.weights_bytes(sub))
    wp.generate(d_delays, epoch, d_chk)
    np.testing.assert_array_equal(d_w.download(), d_chk.download())


@pytest.mark.parametrize("ts", [4, 16, 64])
def test_config3_batch_graph_bit_exact(gpu, oracle, ts):
    capi = gpu
    from psrdada_cpp_amd import workloads as wl
    cfg = wl.config3(ts)
    F, nhg, T = cfg.nchans, cfg.nheap_groups, cfg.nsamples
    print(f"config 3 block: 64 ant split 40/16/8, 768/256/64 beams, {F} ch, {T} spectra, ts {ts}", flush=True)
    subs = [dict(start_antenna=sb.start_antenna, nantennas=sb.nantennas, nbeams=sb.nbeams, tscrunch=ts,
                 fscrunch=1, nsamples_per_heap_out=T // ts) for sb in cfg.subarrays]
    levels = [cfg.levels(sb)[0] for sb in cfg.subarrays]
    batch = capi.Batch(cfg.total_nantennas, F, nhg, subs, levels)
    (ibo, ibs) = cfg.levels(cfg.subarrays[0])[1]
    batch.enable_ib(cfg.ib_tscrunch, cfg.ib_fscrunch, ibo, ibs)
    batch.autotune(reps=2)
    ws = [_weights(capi, wl, cfg, sb, 50 + i) for i, sb in enumerate(cfg.subarrays)]
    outs = [capi.DeviceBuffer(batch.output_bytes(i)) for i in range(len(subs) + 1)]
    s = capi.Stream()
    scratch = capi.DeviceBuffer(1 << 20)
    d_taftp = None
    for rep in range(2):
        taftp = wl.synth_taftp(cfg, 31 + ts + 100 * rep)
        if d_taftp is None:
            d_taftp = capi.DeviceBuffer.from_array(taftp)
        else:
            d_taftp.upload(taftp, s)
        for o in outs:
            o.fill(0x55, s)
        batch.run(d_taftp, [d for d, _ in ws], outs, use_graph=True, stream=s)
        s.sync()
        scratch.fill(rep, s)          # eager work between replays
        s.sync()
        for k, sb in enumerate(cfg.subarrays):
            ftpa = oracle.split_transpose(taftp, nhg, cfg.total_nantennas, sb.nantennas, sb.start_antenna, F)
            p = oracle.cb_power(ftpa, ws[k][1], F, T, sb.nantennas, sb.nbeams, ts, 1, nthreads=16)
            np.testing.assert_array_equal(outs[k].download(), oracle.cb_tbtf(p, levels[k][0], levels[k][1], T // ts),
                                          err_msg=f"replay {rep} subarray {k}")
        pib = oracle.ib_power(taftp, F, cfg.ib_tscrunch, cfg.ib_fscrunch, nhg, cfg.total_nantennas, nthreads=16)
        np.testing.assert_array_equal(outs[-1].download(), oracle.ib_tf(pib, ibo, ibs), err_msg=f"replay {rep} IB")
    batch.close()


def test_config5_serving_feeds_bit_exact(gpu, oracle):
    capi = gpu
    from psrdada_cpp_amd import serving, workloads as wl
    feeds = serving.four_subarray_feeds(32)
    print("config 5: one 32-heap-group block of each feed " +
          ", ".join(f"{f.nantennas}/{f.nbeams}/{f.nchans}" for f in feeds), flush=True)
    srv, fids = serving.make_server(feeds, 16 << 30, nslots=4, seed=3)
    try:
        blocks = []
        for k, (fid, fs) in enumerate(zip(fids, feeds)):
            bb, cbb, ibb = srv.feed_bytes(fid)
            x = capi.HostBuffer(bb)
            x.array[:] = capi.synth_gaussian_int8(900 + k, fs.input_level, bb)
            cb_o, ib_o = capi.HostBuffer(cbb), capi.HostBuffer(ibb)
            srv.submit(fid, x, cb_o, ib_o, tag=k)
            blocks.append((x, cb_o, ib_o))
        for _ in feeds:
            r = srv.poll(timeout_ms=30000)
            assert r is not None and r[1] == 0
    finally:
        srv.close()
    for k, fs in enumerate(feeds):
        A, B, F, T = fs.nantennas, fs.nbeams, fs.nchans, fs.nsamples
        # the weights the engine generated: same delays, frequencies and kernel
        delays = np.zeros((B * A, 2), np.float32)
        delays[:, 1] = capi.synth_uniform_f32(3 + 17 * k, -1e-6, 1e-6, B * A)
        freqs = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F * wl.CHANNEL_RATE_HZ, F)
        wp = capi.WeightsPlan(A, B, F, freqs)
        d_w = capi.DeviceBuffer(F * B * A * 2)
        wp.generate(capi.DeviceBuffer.from_array(delays.reshape(-1)), 0.0, d_w)
        w = d_w.download()
        cbo, cbs, ibo, ibs = capi.power_levels(fs.input_level, fs.output_level, A, fs.tscrunch, fs.fscrunch, A,
                                               fs.ib_tscrunch, fs.ib_fscrunch, 2)
        x, cb_o, ib_o = blocks[k]
        ftpa = oracle.split_transpose(x.array, fs.nheap_groups, A, A, 0, F)
        p = oracle.cb_power(ftpa, w, F, T, A, B, fs.tscrunch, fs.fscrunch, nthreads=16)
        np.testing.assert_array_equal(cb_o.array, oracle.cb_tbtf(p, np.full(F, cbo, np.float32),
                                                                 np.full(F, cbs, np.float32), T // fs.tscrunch),
                                      err_msg=f"feed {k} CB")
        pib = oracle.ib_power(x.array, F, fs.ib_tscrunch, fs.ib_fscrunch, fs.nheap_groups, A, nthreads=16)
        np.testing.assert_array_equal(ib_o.array, oracle.ib_tf(pib, np.full(F, ibo, np.float32),
                                                               np.full(F, ibs, np.float32)), err_msg=f"feed {k} IB")
