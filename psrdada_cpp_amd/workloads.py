"""Standard FBFUSE beamformer workloads and their seeded synthetic inputs.

Geometry of one DADA block (TAFTP) and the beamformer settings for the
configurations the project is measured on.  Inputs are generated natively
(``fbf_synth_*``: splitmix64 + Box-Muller, documented in capi.cpp) so that the
same seed gives the same bytes in Python and in C++.

Constants follow fbfuse_constants.hpp:6-8 (MeerKAT F-engine: 4096 channels,
256 samples per heap, 2 pols) and Pipeline.cu:91 (2 * 4096 sample-clock ticks
per spectrum at 1712 MHz, i.e. a channel rate of 208 984.375 Hz).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import capi

SAMPLE_CLOCK_HZ = 1712e6
NCHANS_TOTAL = 4096
CHANNEL_RATE_HZ = SAMPLE_CLOCK_HZ / (2 * NCHANS_TOTAL)  # 208 984.375 Hz
CHANNEL_BW_HZ = CHANNEL_RATE_HZ
NSAMPLES_PER_HEAP = 256
NPOL = 2
L_BAND_CENTRE_HZ = 1284e6

# MI355X peaks used for roofline fractions (vendor figures, dense).
PEAK_HBM_GBS = 8000.0
PEAK_INT8_TOPS = 5000.0


@dataclass
class Subarray:
    """One coherent-beam product: a contiguous antenna subset and its beams."""
    nantennas: int
    nbeams: int
    start_antenna: int = 0
    tscrunch: int = 16
    fscrunch: int = 1


@dataclass
class BlockConfig:
    name: str
    total_nantennas: int
    nchans: int
    nheap_groups: int
    subarrays: list = field(default_factory=list)
    ib_tscrunch: int = 16
    ib_fscrunch: int = 1
    input_level: float = 16.0      # sigma of re/im of the voltages
    output_level: float = 24.0     # PipelineConfig.cpp:19 default
    cfreq_hz: float = L_BAND_CENTRE_HZ
    bw_hz: float = 0.0             # 0 -> nchans * CHANNEL_BW_HZ
    cb_heap_bytes: int = 8192      # FBFUSE_CB_HEAP_SIZE
    tau_max_s: float = 1e-6        # weights: tau ~ U(-tau_max, tau_max)

    @property
    def nsamples(self):
        return self.nheap_groups * NSAMPLES_PER_HEAP

    @property
    def bandwidth_hz(self):
        return self.bw_hz if self.bw_hz else self.nchans * CHANNEL_BW_HZ

    @property
    def taftp_bytes(self):
        return self.nheap_groups * self.total_nantennas * self.nchans * NSAMPLES_PER_HEAP * NPOL * 2

    @property
    def block_seconds(self):
        return self.nsamples / CHANNEL_RATE_HZ

    def ftpa_bytes(self, sub):
        return self.nchans * self.nsamples * NPOL * sub.nantennas * 2

    def weights_bytes(self, sub):
        return self.nchans * sub.nbeams * sub.nantennas * 2

    def cb_out_bytes(self, sub):
        return (self.nsamples // sub.tscrunch) * sub.nbeams * (self.nchans // sub.fscrunch)

    def cb_ohs(self, sub):
        return self.cb_heap_bytes // (self.nchans // sub.fscrunch)

    def ib_out_bytes(self):
        return (self.nsamples // self.ib_tscrunch) * (self.nchans // self.ib_fscrunch)

    def cb_ops(self, sub=None):
        """Algorithmic int8 ops per block: 8 * F * T * P * A * B (complex MAC = 8 ops)."""
        subs = [sub] if sub is not None else self.subarrays
        return sum(8 * self.nchans * self.nsamples * NPOL * s.nantennas * s.nbeams for s in subs)

    def levels(self, sub):
        """Per-output-channel (offsets, scales) for CB of `sub` and for the IB
        (PipelineConfig.cpp:109-127 evaluated natively; equal for every channel
        because the synthetic voltages have the same level in each)."""
        cbo, cbs, ibo, ibs = capi.power_levels(self.input_level, self.output_level, sub.nantennas,
                                               sub.tscrunch, sub.fscrunch, self.total_nantennas,
                                               self.ib_tscrunch, self.ib_fscrunch, NPOL)
        fo = self.nchans // sub.fscrunch
        fib = self.nchans // self.ib_fscrunch
        return ((np.full(fo, cbo, np.float32), np.full(fo, cbs, np.float32)),
                (np.full(fib, ibo, np.float32), np.full(fib, ibs, np.float32)))

    def channel_frequencies(self):
        return capi.channel_frequencies_f64(self.cfreq_hz, self.bandwidth_hz, self.nchans)


def config1():
    """CPU host-reference plumbing case: 16 ant, 16 ch, 8 heap groups, 32 beams, ts 16.
    Channel centres span 856-1712 MHz (53.5 MHz channels); output heap of 2048 B
    (128 samples x 16 channels) so that one block fills whole heaps."""
    return BlockConfig("config1_host_ref_16ant_16ch_32beams", total_nantennas=16, nchans=16, nheap_groups=8,
                       subarrays=[Subarray(16, 32)], bw_hz=856e6, cb_heap_bytes=2048)


def config2():
    """Single-subarray GPU beamformer: 64 ant, 64 ch, 32 heap groups, 512 beams, ts 16, + IB."""
    return BlockConfig("config2_single_subarray_64ant_64ch_512beams", total_nantennas=64, nchans=64,
                       nheap_groups=32, subarrays=[Subarray(64, 512)])


def config3(tscrunch=16):
    """Split + batched: one 64-antenna block split 40/16/8 with 768/256/64 beams."""
    return BlockConfig(f"config3_split_40_16_8_ts{tscrunch}", total_nantennas=64, nchans=64, nheap_groups=32,
                       subarrays=[Subarray(40, 768, 0, tscrunch), Subarray(16, 256, 40, tscrunch),
                                  Subarray(8, 64, 56, tscrunch)])


# ---------------------------------------------------------------------------
# Synthetic inputs
# ---------------------------------------------------------------------------
def synth_taftp(cfg, seed):
    """TAFTP voltages: re/im = clamp(round(N(0, input_level)), -127, 127)."""
    return capi.synth_gaussian_int8(seed, cfg.input_level, cfg.taftp_bytes, clip=True)


def synth_delays(nbeams, nantennas, seed, tau_max_s, rate_max=0.0):
    """Delay models [B][A] float2: x = rate ~ U(-rate_max, rate_max), y = offset ~ U(-tau_max, tau_max)."""
    n = nbeams * nantennas
    d = np.empty((n, 2), dtype=np.float32)
    d[:, 1] = capi.synth_uniform_f32(seed, -tau_max_s, tau_max_s, n)
    if rate_max:
        d[:, 0] = capi.synth_uniform_f32(seed ^ 0x5DEECE66D, -rate_max, rate_max, n)
    else:
        d[:, 0] = 0.0
    return d.reshape(-1)
