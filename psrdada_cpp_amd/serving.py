"""Serving workload driver: several subarray feeds of uneven antenna and beam
counts, blocks arriving as Poisson streams, served by the native engine
(``fbf_server_*``: continuous batching over a pool of HIP streams, weights
device-resident, admission bounded by an HBM budget).

A feed's real-time rate is one block per block duration (``nheap_groups`` x
256 spectra at 208 984.375 spectra/s: 39.2 ms for 32 heap groups), so an
aggregate load of ``x`` real time is a Poisson stream of ``x / block_s``
blocks/s per feed.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np

from . import capi
from .workloads import CHANNEL_RATE_HZ, L_BAND_CENTRE_HZ, NPOL, NSAMPLES_PER_HEAP


@dataclass
class FeedSpec:
    nantennas: int
    nbeams: int
    nchans: int
    nheap_groups: int = 32
    tscrunch: int = 16
    fscrunch: int = 1
    ib_tscrunch: int = 16
    ib_fscrunch: int = 1
    input_level: float = 16.0
    output_level: float = 24.0

    @property
    def nsamples(self):
        return self.nheap_groups * NSAMPLES_PER_HEAP

    @property
    def block_seconds(self):
        return self.nsamples / CHANNEL_RATE_HZ

    @property
    def block_bytes(self):
        return self.nheap_groups * self.nantennas * self.nchans * NSAMPLES_PER_HEAP * NPOL * 2

    @property
    def cb_ops(self):
        return 8 * self.nchans * self.nsamples * NPOL * self.nantennas * self.nbeams


def four_subarray_feeds(nheap_groups=32):
    """64 ant/768 beams/64 ch, 32/256/64, 16/128/32, 8/64/16."""
    return [FeedSpec(64, 768, 64, nheap_groups), FeedSpec(32, 256, 64, nheap_groups),
            FeedSpec(16, 128, 32, nheap_groups), FeedSpec(8, 64, 16, nheap_groups)]


def make_server(feeds, hbm_budget, nslots=8, seed=1, tau_max_s=1e-6):
    """Creates the engine, adds the feeds with delay models (weights generated
    on the device) and starts it.  Returns (server, feed ids)."""
    srv = capi.Server(nslots, hbm_budget)
    fids = []
    for k, fs in enumerate(feeds):
        cbo, cbs, ibo, ibs = capi.power_levels(fs.input_level, fs.output_level, fs.nantennas, fs.tscrunch,
                                               fs.fscrunch, fs.nantennas, fs.ib_tscrunch, fs.ib_fscrunch, NPOL)
        fo, fib = fs.nchans // fs.fscrunch, fs.nchans // fs.ib_fscrunch
        fid = srv.add_feed(fs.nantennas, fs.nbeams, fs.nchans, fs.nheap_groups, fs.tscrunch, fs.fscrunch,
                           (np.full(fo, cbo, np.float32), np.full(fo, cbs, np.float32)),
                           ib=(fs.ib_tscrunch, fs.ib_fscrunch),
                           ib_levels=(np.full(fib, ibo, np.float32), np.full(fib, ibs, np.float32)),
                           nsamples_per_heap_out=fs.nsamples // fs.tscrunch)
        delays = np.zeros((fs.nbeams * fs.nantennas, 2), np.float32)
        delays[:, 1] = capi.synth_uniform_f32(seed + 17 * k, -tau_max_s, tau_max_s, fs.nbeams * fs.nantennas)
        bw = fs.nchans * CHANNEL_RATE_HZ
        srv.set_delays(fid, delays.reshape(-1), capi.channel_frequencies_f64(L_BAND_CENTRE_HZ, bw, fs.nchans), 0.0)
        fids.append(fid)
    srv.start()
    return srv, fids


class _Buffers:
    """Per-feed pinned input blocks (synthetic, reused read-only) and a pool of
    output buffers handed out one per block in flight."""

    def __init__(self, srv, fids, feeds, seed, nin, nout, device_input):
        self.inputs, self.free_out, self.sem = [], [], []
        for k, (fid, fs) in enumerate(zip(fids, feeds)):
            bb, cbb, ibb = srv.feed_bytes(fid)
            ins = []
            for i in range(nin):
                x = capi.synth_gaussian_int8(seed + 1000 * k + i, fs.input_level, bb)
                if device_input:
                    ins.append(capi.DeviceBuffer.from_array(x))
                else:
                    h = capi.HostBuffer(bb)
                    h.array[:] = x
                    ins.append(h)
            self.inputs.append(ins)
            self.free_out.append([(capi.HostBuffer(cbb), capi.HostBuffer(ibb)) for _ in range(nout)])
            self.sem.append(threading.Semaphore(nout))
        self.lock = threading.Lock()

    def take(self, k):
        self.sem[k].acquire()
        with self.lock:
            return self.free_out[k].pop()

    def give(self, k, o):
        with self.lock:
            self.free_out[k].append(o)
        self.sem[k].release()


def _pct(x):
    x = np.asarray(x) if len(x) else np.zeros(1)
    return {"p50": round(float(np.percentile(x, 50)), 3), "p99": round(float(np.percentile(x, 99)), 3),
            "max": round(float(x.max()), 3)}


def _summary(srv, feeds, counts, wall_s, block_lat, feed_lat=None):
    st = srv.stats()
    data_s = sum(c * fs.block_seconds for c, fs in zip(counts, feeds))
    ops = sum(c * fs.cb_ops for c, fs in zip(counts, feeds))
    nbytes = sum(c * fs.block_bytes for c, fs in zip(counts, feeds))
    lat = np.asarray(block_lat) if block_lat else np.zeros(1)
    per_feed = {}
    if feed_lat is not None:
        per_feed = {f"{fs.nantennas}/{fs.nbeams}/{fs.nchans}": _pct(fl) for fs, fl in zip(feeds, feed_lat)}
    return {
        "blocks": int(sum(counts)),
        "blocks_per_feed": [int(c) for c in counts],
        "wall_s": round(wall_s, 3),
        # data seconds served per wall second, per feed: the factor by which
        # all feeds together run ahead of real time
        "x_real_time": round(data_s / len(feeds) / wall_s, 3),
        "cb_int8_tops": round(ops / wall_s / 1e12, 2),
        "input_gbs": round(nbytes / wall_s / 1e9, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3),
                       "mean": round(float(lat.mean()), 3)},
        "latency_ms_per_feed": per_feed,
        "engine_p99_ms": round(st["p99_ms"], 3),
        "failed": int(st["failed"]),
        "hbm_budget": int(st["hbm_budget"]),
        "hbm_peak": int(st["hbm_peak"]),
    }


def _collector(srv, bufs, owner, n, lat, counts, errors, feed_lat=None):
    """Polls n completions (n None: until the owner map holds a None entry
    and is otherwise empty), returns their output buffers to the pool."""
    got = 0
    last = time.perf_counter()
    try:
        while n is None or got < n:
            if n is None and owner.get(-1, 0) is None and len(owner) == 1:
                return
            # Short polls when the count is open, so that the end marker is
            # seen; 30 s without a completion is a failure either way.
            r = srv.poll(timeout_ms=30000 if n is not None else 50)
            if r is None:
                if n is None and time.perf_counter() - last < 30.0:
                    continue
                errors.append("no completion within 30 s")
                return
            last = time.perf_counter()
            tag, status, ms = r
            k, o = owner.pop(tag)
            bufs.give(k, o)
            if status != 0:
                errors.append(f"block {tag} failed with status {status}")
            else:
                lat.append(ms)
                counts[k] += 1
                if feed_lat is not None:
                    feed_lat[k].append(ms)
            got += 1
    except Exception as e:  # reported by the caller after join()
        errors.append(repr(e))


def warm_up(srv, fids, feeds, seed=3, device_input=False):
    """One untimed block of every feed, so that code-object loading and
    first-launch costs are not counted as serving latency."""
    bufs = _Buffers(srv, fids, feeds, seed, 1, 1, device_input)
    # The output buffers of every block in flight stay referenced (the
    # engine copies into them) until its completion has been polled.
    outs = []
    for k, fid in enumerate(fids):
        outs.append(bufs.take(k))
        srv.submit(fid, bufs.inputs[k][0], outs[k][0], outs[k][1], tag=10**9 + k, on_device=device_input)
    for _ in fids:
        r = srv.poll(timeout_ms=30000)
        if r is None or r[1] != 0:
            # an unfinished block may still write its outputs: wait for all
            srv.drain()
            raise RuntimeError(f"warm-up block failed: {r}")
    del outs


def run_poisson(srv, fids, feeds, load_x=1.5, duration_s=10.0, seed=7, device_input=False, nout=8):
    """Poisson arrivals at `load_x` real time per feed for `duration_s`
    seconds; returns throughput and latency figures."""
    rng = np.random.default_rng(seed)
    arrivals = []
    for k, fs in enumerate(feeds):
        rate = load_x / fs.block_seconds
        t = rng.exponential(1.0 / rate)
        while t < duration_s:
            arrivals.append((t, k))
            t += rng.exponential(1.0 / rate)
    arrivals.sort()
    warm_up(srv, fids, feeds, seed, device_input)
    bufs = _Buffers(srv, fids, feeds, seed, 2, nout, device_input)
    owner, lat, counts, errors = {}, [], [0] * len(feeds), []
    feed_lat = [[] for _ in feeds]
    srv.reset_stats()
    col = threading.Thread(target=_collector, args=(srv, bufs, owner, len(arrivals), lat, counts, errors, feed_lat))
    col.start()
    t0 = time.perf_counter()
    late = 0
    for tag, (ta, k) in enumerate(arrivals):
        wait = ta - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        elif wait < -0.005:
            late += 1
        o = bufs.take(k)
        owner[tag] = (k, o)
        srv.submit(fids[k], bufs.inputs[k][tag % 2], o[0], o[1], tag=tag, on_device=device_input)
    col.join()
    wall = time.perf_counter() - t0
    if errors:
        srv.drain()   # blocks still running write into the pools' buffers
        raise RuntimeError("; ".join(errors))
    out = _summary(srv, feeds, counts, wall, lat, feed_lat)
    out.update({"offered_load_x": load_x, "arrivals": "poisson", "late_submissions": late,
                "input": "device" if device_input else "pinned host (PCIe-inclusive)"})
    return out


def run_saturation(srv, fids, feeds, nblocks_per_feed=50, device_input=False, seed=11, nout=4, seconds=None):
    """Closed loop: blocks of all feeds round-robin, at most `nout` outstanding
    per feed, submitted as fast as the engine completes them; either
    `nblocks_per_feed` blocks of each feed or, with `seconds`, for that long
    (after one untimed warm-up block per feed)."""
    warm_up(srv, fids, feeds, seed, device_input)
    bufs = _Buffers(srv, fids, feeds, seed, 2, nout, device_input)
    owner, lat, counts, errors = {}, [], [0] * len(feeds), []
    feed_lat = [[] for _ in feeds]
    total = None if seconds else nblocks_per_feed * len(feeds)
    srv.reset_stats()
    col = threading.Thread(target=_collector, args=(srv, bufs, owner, total, lat, counts, errors, feed_lat))
    t0 = time.perf_counter()
    col.start()
    tag = 0
    while (tag < total) if total is not None else (time.perf_counter() - t0 < seconds):
        k = tag % len(feeds)
        o = bufs.take(k)
        owner[tag] = (k, o)
        srv.submit(fids[k], bufs.inputs[k][tag % 2], o[0], o[1], tag=tag, on_device=device_input)
        tag += 1
    if total is None:
        owner[-1] = None   # no more submissions: the collector ends once the rest completed
    col.join()
    wall = time.perf_counter() - t0
    if errors:
        srv.drain()   # blocks still running write into the pools' buffers
        raise RuntimeError("; ".join(errors))
    out = _summary(srv, feeds, counts, wall, lat, feed_lat)
    out.update({"arrivals": "closed loop", "outstanding_per_feed": nout,
                "input": "device" if device_input else "pinned host (PCIe-inclusive)"})
    return out
