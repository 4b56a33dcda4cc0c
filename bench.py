#!/usr/bin/env python3
"""FBFUSE beamformer benchmark on one MI355X.

A step is one DADA block of the single-subarray workload (64 antennas,
64 channels, 32 heap groups = 8192 spectra = 39.199 ms of data, 512 coherent
beams, tscrunch 16, plus the incoherent beam), with the voltages already
resident in HBM (three distinct blocks, 403 MB together, rotate so that the
input is not served from the 256 MB Infinity Cache):
    coherent beamformer straight from TAFTP (staged and transposed in LDS,
        no split-transpose pass; int8 MFMA, detect, scrunch, quantise)
        -> TBTF int8, the same kernel leaving the incoherent beam's
        per-sample partials
    TBTF reorder, whose launch also finishes the incoherent beam (TF int8)
        and generates the next block's weights (delay -> weights at that
        block's epoch, FP64 phase: every block's weights are generated
        once, as Pipeline.cu:164 does at the block's start)
all on one stream (--path taftp: weights + incoherent beam in one launch,
then the coherent beamformer from TAFTP; --path ftpa: the reference's
order, a split-transpose pass to FTPA with the incoherent beam and the
weights in it, then the coherent beamformer on FTPA).  Prints one JSON line.

value        = x real-time: block duration / wall time per block (K blocks
               timed back to back, device synchronised on both sides).
roofline     = the coherent-beam kernel: algorithmic int8 ops per launch
               (8*F*T*P*A*B) / its mean duration (HIP events on its stream)
               vs the 5.0 POP/s dense int8 MFMA peak; practical_reference
               is the vendor int8 GEMM on random operands in the same run
               (the int8 MFMA rate depends on the operand data).
cpu_baseline = the CPU oracle (this repo's C restatement, OpenMP) on a bounded
               sample of the same block, scaled to x real-time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-channels", type=int, default=64, help="channels in the CPU sample (64 = whole block)")
    ap.add_argument("--variant", type=int, default=0, help="CB kernel variant (0 = auto)")
    ap.add_argument("--path", choices=["fused", "taftp", "ftpa"], default="fused",
                    help="fused: weights, then coherent and incoherent beams in one pass over TAFTP; "
                         "taftp: weights + IB in one launch, then CB straight from TAFTP; "
                         "ftpa: split-transpose pass (with IB and weights), then CB on FTPA")
    ap.add_argument("--separate-ib", action="store_true",
                    help="run the incoherent beam as its own kernel instead of inside the split-transpose pass")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 40/16/8 subarray leg")
    ap.add_argument("--no-corner-turns", action="store_true",
                    help="skip the TUSE corner-turn and F-engine reorder leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the four-feed serving leg")
    ap.add_argument("--no-mfma-reference", action="store_true",
                    help="skip the vendor int8 GEMM rate measured beside the roofline")
    ap.add_argument("--mfma-reference-only", action="store_true",
                    help="print the vendor int8 GEMM rate as one JSON line and exit (run as a child process)")
    ap.add_argument("--serving-seconds", type=float, default=8.0, help="Poisson arrival window of the serving leg")
    ap.add_argument("--hbm-budget-gb", type=float, default=64.0, help="serving engine HBM budget")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 3 times (scale for the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--no-extra-legs", action="store_true",
                    help="skip the Pipeline, weights-generation and config-1 legs")
    ap.add_argument("--profile-leg", choices=["batch", "corner_turns", "serving", "pipeline", "weights", "config1"],
                    help="run only this leg and print its JSON (for profiling one workload)")
    return ap.parse_args()


SETTLE_S = 0.5


def percentile(x, q):
    return float(np.percentile(np.asarray(x), q))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; --gpus above 1 is not supported "
                                   "(the path shards by frequency channel with no exchange between GPUs)"}))
        return 2

    if args.mfma_reference_only:
        print(json.dumps(mfma_reference()), flush=True)
        return 0
    if args.calibrate_copy:
        return calibrate_copy()
    if args.profile_leg:
        return profile_leg(args)

    from psrdada_cpp_amd import capi, workloads as wl
    import ctypes

    n = ctypes.c_int()
    capi.call("fbf_device_count", ctypes.byref(n))
    if n.value < 1:
        raise RuntimeError("no GPU visible")
    capi.call("fbf_set_device", 0)
    name = ctypes.create_string_buffer(256)
    capi.call("fbf_device_name", name, 256)

    cfg = wl.config2()
    sub = cfg.subarrays[0]
    A, F, B, T = sub.nantennas, cfg.nchans, sub.nbeams, cfg.nsamples
    seed = 20260101
    NBLK = 3
    taftp = wl.synth_taftp(cfg, seed)
    delays = wl.synth_delays(B, A, seed + 1, cfg.tau_max_s)
    freqs = cfg.channel_frequencies()
    (cb_off, cb_sc), (ib_off, ib_sc) = cfg.levels(sub)

    s = capi.Stream()
    d_taftps = [capi.DeviceBuffer.from_array(taftp if i == 0 else wl.synth_taftp(cfg, seed + 100 + i), s)
                for i in range(NBLK)]
    fused_path = args.path == "fused"
    taftp_path = args.path in ("taftp", "fused")
    d_ftpa = None if taftp_path else capi.DeviceBuffer(cfg.ftpa_bytes(sub))
    d_delays = capi.DeviceBuffer.from_array(delays, s)
    d_w = capi.DeviceBuffer(cfg.weights_bytes(sub))
    wp = capi.WeightsPlan(A, B, F, freqs)
    cb = capi.CbPlan(A, B, F, sub.tscrunch, sub.fscrunch, cb_off, cb_sc,
                     nsamples_per_heap_out=cfg.cb_ohs(sub))
    if args.variant:
        cb.set_variant(args.variant)
    ib = capi.IbPlan(cfg.total_nantennas, F, cfg.ib_tscrunch, cfg.ib_fscrunch, ib_off, ib_sc)
    d_cb = capi.DeviceBuffer(cfg.cb_out_bytes(sub))
    d_ib = capi.DeviceBuffer(cfg.ib_out_bytes())

    # FTPA path: the incoherent beam comes out of the split-transpose pass
    # when the block is split whole and the IB scrunch allows it
    # (fbf_split_transpose_ib).
    fused = (not taftp_path and A == cfg.total_nantennas and cfg.ib_fscrunch == 1 and 32 % cfg.ib_tscrunch == 0
             and not args.separate_ib)
    if taftp_path and not cb.taftp_direct(cfg.total_nantennas, T):
        raise RuntimeError("the coherent-beam plan does not read this block straight from TAFTP")
    if fused_path and not cb.ib_fused(ib, T):
        raise RuntimeError("the plans do not form the incoherent beam inside the coherent kernel")
    # Epoch of the weights last written to d_w (for the CPU baseline's check).
    w_epoch = [0.0]
    if fused_path:
        # Block k's weights come from block k - 1's reorder launch; the first
        # block's are generated here.
        wp.generate(d_delays, 0.0, d_w, stream=s)

    def step(k, ev=None, weights_every=1):
        """One block on stream s.  A block that takes new weights generates
        them in the incoherent beam's launch (fbf_ib_beamform_weights; FTPA
        path: in the split-transpose's, fbf_split_transpose_ib_weights);
        with `ev` (five events) every kernel runs on its own, in turn, for
        per-kernel durations."""
        d_taftp = d_taftps[k % NBLK]
        new_w = k % weights_every == 0
        epoch = cfg.block_seconds * k
        if ev:
            ev[0].record(s)
            if new_w:
                wp.generate(d_delays, epoch, d_w, stream=s)
                w_epoch[0] = epoch
            ev[1].record(s)
            if fused_path:
                pass
            elif taftp_path:
                ib.beamform(d_taftp, d_ib, stream=s)
            elif fused:
                capi.split_transpose_ib(d_taftp, d_ftpa, F, ib, d_ib, stream=s)
            else:
                capi.split_transpose(d_taftp, d_ftpa, cfg.total_nantennas, A, 0, F, stream=s)
            ev[2].record(s)
        elif fused_path:
            pass
        elif taftp_path:
            if new_w:
                ib.beamform_weights(d_taftp, d_ib, wp, d_delays, epoch, d_w, stream=s)
                w_epoch[0] = epoch
            else:
                ib.beamform(d_taftp, d_ib, stream=s)
        elif fused and new_w:
            capi.split_transpose_ib_weights(d_taftp, d_ftpa, F, ib, d_ib, wp, d_delays, epoch, d_w, stream=s)
            w_epoch[0] = epoch
        else:
            if new_w:
                wp.generate(d_delays, epoch, d_w, stream=s)
                w_epoch[0] = epoch
            if fused:
                capi.split_transpose_ib(d_taftp, d_ftpa, F, ib, d_ib, stream=s)
            else:
                capi.split_transpose(d_taftp, d_ftpa, cfg.total_nantennas, A, 0, F, stream=s)
        if fused_path:
            # The reorder's launch generates the next block's weights when
            # that block takes new ones.
            if (k + 1) % weights_every == 0:
                cb.ib_beamform_taftp_next(ib, d_taftp, d_w, d_cb, d_ib, wp, d_delays, epoch + cfg.block_seconds,
                                          d_w, stream=s)
                w_epoch[0] = epoch + cfg.block_seconds
            else:
                cb.ib_beamform_taftp(ib, d_taftp, d_w, d_cb, d_ib, stream=s)
        elif taftp_path:
            cb.beamform_taftp(d_taftp, d_w, d_cb, nantennas_total=cfg.total_nantennas, stream=s)
        else:
            cb.beamform(d_ftpa, d_w, d_cb, stream=s)
        if ev:
            ev[3].record(s)
        if not fused and not taftp_path:
            ib.beamform(d_taftp, d_ib, stream=s)
        if ev:
            ev[4].record(s)

    def timed_blocks(K, k0, weights_every, kernel_events=False):
        """K blocks back to back, events at block boundaries (and, with
        kernel_events, the plan's events right before and after the coherent
        kernel of every block: fbf_cb_plan_set_timing)."""
        evs = [capi.Event() for _ in range(K + 1)]
        kev = [(capi.Event(), capi.Event()) for _ in range(K)] if kernel_events else None
        capi.call("fbf_device_synchronize")
        t0 = time.perf_counter()
        evs[0].record(s)
        for k in range(K):
            if kev:
                cb.set_timing(*kev[k])
            step(k0 + k, weights_every=weights_every)
            evs[k + 1].record(s)
        s.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3
        if kev:
            cb.set_timing(None, None)
        kern = [b.elapsed_ms(a) for b, a in kev] if kev else None
        return wall_ms, [evs[k].elapsed_ms(evs[k + 1]) for k in range(K)], evs[0].elapsed_ms(evs[K]), kern

    # Clock settling: the GPU idles at a low clock and its power management
    # ramps up over tens of milliseconds of load, so untimed steps of the
    # same workload run (25 at a time, synchronised) until it has been busy
    # for SETTLE_S; a beamformer that runs continuously sees the settled
    # clock.  Then the W warmup steps and exactly K timed steps.
    t_settle = time.perf_counter()
    n_settle = 0
    while time.perf_counter() - t_settle < SETTLE_S:
        for _ in range(25):
            step(n_settle)
            n_settle += 1
        s.sync()
    settle_s = time.perf_counter() - t_settle
    for k in range(args.warmup):
        step(k)
    s.sync()

    # Timed region: exactly K blocks, weights regenerated every block.
    K = args.steps
    wall_ms, block_ms, dev_total_ms, kern_ms = timed_blocks(K, args.warmup, 1, kernel_events=True)
    ms_per_step = wall_ms / K

    # The same K blocks with the weights refreshed every 8th block (the
    # delay models change slowly; the serving engine refreshes on demand).
    wall8_ms, _, _, _ = timed_blocks(K, args.warmup + K, 8)

    # Per-kernel durations (separate pass, events around every kernel).
    R = min(50, max(10, K // 4))
    kev = [[capi.Event() for _ in range(5)] for _ in range(R)]
    for r in range(R):
        step(args.warmup + 2 * K + r, kev[r])
    s.sync()
    # Medians: the pass starts from an idle queue, so the first blocks can
    # include host-side launch gaps.
    w_ms = np.median([kev[r][0].elapsed_ms(kev[r][1]) for r in range(R)])
    st_ms = np.median([kev[r][1].elapsed_ms(kev[r][2]) for r in range(R)])
    cb_ms = np.median([kev[r][2].elapsed_ms(kev[r][3]) for r in range(R)])
    ib_ms = np.median([kev[r][3].elapsed_ms(kev[r][4]) for r in range(R)])
    # The launch the timed blocks use for a block's weights: IB + weights
    # (FTPA path: split-transpose + IB + weights).
    stw_ms = None
    if fused or (taftp_path and not fused_path):
        fe = [capi.Event() for _ in range(R + 1)]
        fe[0].record(s)
        for r in range(R):
            w_epoch[0] = cfg.block_seconds * r
            if taftp_path:
                ib.beamform_weights(d_taftps[r % NBLK], d_ib, wp, d_delays, cfg.block_seconds * r, d_w, stream=s)
            else:
                capi.split_transpose_ib_weights(d_taftps[r % NBLK], d_ftpa, F, ib, d_ib, wp, d_delays,
                                                cfg.block_seconds * r, d_w, stream=s)
            fe[r + 1].record(s)
        s.sync()
        stw_ms = np.median([fe[r].elapsed_ms(fe[r + 1]) for r in range(R)])

    ops = cfg.cb_ops()
    block_s = cfg.block_seconds
    rt = block_s * 1e3 / ms_per_step
    # The coherent kernel alone, from its events in the timed blocks; the
    # window (the coherent kernel and what follows it in the call: the
    # reorder, and on the one-pass path the IB finish and the next weights)
    # from the per-kernel pass.
    cb_kernel_ms = float(np.mean(kern_ms))
    cb_tops = ops / (cb_kernel_ms * 1e-3) / 1e12
    window_tops = ops / (cb_ms * 1e-3) / 1e12
    pipeline_tops = ops / (ms_per_step * 1e-3) / 1e12
    samples_per_s = (F * T * wl.NPOL * cfg.total_nantennas) / (ms_per_step * 1e-3)
    ib_bytes = cfg.taftp_bytes + cfg.ib_out_bytes()
    if fused_path:
        st_bytes = 0
        st_name = None
    elif taftp_path:
        st_bytes = ib_bytes
        st_name = "incoherent_beamformer"
    else:
        st_bytes = 2 * cfg.ftpa_bytes(sub) + (cfg.ib_out_bytes() if fused else 0)
        st_name = "split_transpose_with_incoherent_beam" if fused else "split_transpose"
    kernel_ms = {"weights": round(float(w_ms), 4),
                 ("coherent_and_incoherent_beams" if fused_path else "coherent_beamformer"): round(float(cb_ms), 4)}
    if st_name:
        kernel_ms[st_name] = round(float(st_ms), 4)
    if stw_ms is not None:
        kernel_ms["ib_weights_one_launch" if taftp_path else "split_transpose_ib_weights_one_launch"] = \
            round(float(stw_ms), 4)
    kernel_gbs = {st_name: round(st_bytes / (st_ms * 1e-3) / 1e9, 1)} if st_name else {}
    if not fused and not taftp_path:
        kernel_ms["incoherent_beamformer"] = round(float(ib_ms), 4)
        kernel_gbs["incoherent_beamformer"] = round(ib_bytes / (ib_ms * 1e-3) / 1e9, 1)

    result = {
        "metric": "x real-time & int8 TOPS (64 ant, 64 ch, 512 beams); p99 block latency",
        "value": round(rt, 2),
        "unit": "x real-time",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 (int32 accumulate, exact int64 detect)",
        "data": "synthetic: TAFTP voltages round(N(0,16)) clipped to +-127, seeded splitmix64, 3 distinct "
                "blocks (403 MB) rotating; weights from tau ~ U(-1us, 1us) via the FP64-phase generator, "
                "regenerated every block",
        "config": {
            "workload": cfg.name,
            "nantennas": A, "nchans": F, "nbeams": B, "npol": 2,
            "nsamples_per_block": T, "heap_groups": cfg.nheap_groups,
            "cb_tscrunch": sub.tscrunch, "cb_fscrunch": sub.fscrunch,
            "ib_tscrunch": cfg.ib_tscrunch, "ib_fscrunch": cfg.ib_fscrunch,
            "block_ms": round(block_s * 1e3, 3),
            "stages": ("the coherent and incoherent beams in one pass over TAFTP (the coherent kernel leaves "
                       "per-sample IB partials; no split-transpose pass), then the TBTF reorder, whose launch also "
                       "finishes the IB and generates the next block's weights (every block's weights once)"
                       if fused_path
                       else "weights (every block) and the incoherent beam in one launch, then the coherent beams "
                       "straight from TAFTP (no split-transpose pass)" if taftp_path
                       else "weights (every block), split-transpose and IB in one launch, then CB" if fused
                       else "weights (every block) + split-transpose + CB + IB"),
            "input_blocks_rotating": NBLK,
            "device": name.value.decode(),
        },
        "pipeline_int8_tops": round(pipeline_tops, 2),
        "cb_kernel_int8_tops": round(cb_tops, 2),
        "cb_kernel_ms_timed_blocks": {"mean": round(cb_kernel_ms, 4), "median": round(float(np.median(kern_ms)), 4)},
        "window_int8_tops": round(window_tops, 2),
        "samples_per_s": float(f"{samples_per_s:.4g}"),
        "block_latency_ms": {"p50": round(percentile(block_ms, 50), 4),
                             "p99": round(percentile(block_ms, 99), 4),
                             "max": round(max(block_ms), 4)},
        "device_ms_per_step": round(dev_total_ms / K, 4),
        "clock_settle": {"seconds": round(settle_s, 3), "untimed_steps": n_settle,
                         "why": "the GPU idles at a low clock; untimed steps of the same workload until it has "
                                "been busy for 0.5 s, before the warmup steps"},
        "weights_every_8_blocks": {"ms_per_step": round(wall8_ms / K, 4),
                                   "x_real_time": round(block_s * 1e3 / (wall8_ms / K), 2)},
        "kernel_ms": kernel_ms,
        "kernel_hbm_gbs": kernel_gbs,
        "roofline": {
            "kernel": ("cb_tr_k: the coherent beamformer straight from TAFTP (int8 MFMA, exact detection, scrunch, "
                       "quantise), leaving the incoherent beam's partials" if fused_path else
                       "cb_tr_k: the coherent beamformer straight from TAFTP (int8 MFMA, exact detection, scrunch, "
                       "quantise)" if taftp_path else
                       "cb_tr_k: the coherent beamformer on FTPA (int8 MFMA, exact detection, scrunch, quantise)"),
            "bound": "mfma",
            "achieved": round(cb_tops, 2),
            "peak": wl.PEAK_INT8_TOPS,
            "unit": "TOP/s",
            "frac": round(cb_tops / wl.PEAK_INT8_TOPS, 4),
            "traffic": profiled_traffic(args.path),
            "traffic_source": ("profile constant: rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of cb_tr_k on this "
                               "workload, scaled by a plain copy's readings, from "
                               + os.path.relpath(TRAFFIC_PROFILES[args.path], ROOT) + " (not measured in this run)"
                               if args.path in TRAFFIC_PROFILES else "no committed profile for this path"),
            "ops_per_launch": ops,
            "timing": "HIP events recorded on the beamformer's stream right before and after cb_tr_k's launch in "
                      "every timed block (fbf_cb_plan_set_timing); achieved = ops_per_launch / their mean",
            "window": ("the call around cb_tr_k, from the per-kernel pass: fbf_cb_ib_beamform_taftp_next (cb_tr_k, "
                       "then tbtf_k's launch: the reorder, the IB finish and the next block's weights)" if fused_path
                       else "the call around cb_tr_k, from the per-kernel pass: fbf_cb_beamform_taftp (cb_tr_k, "
                       "then the tbtf_k reorder)" if taftp_path else
                       "the call around cb_tr_k, from the per-kernel pass: fbf_cb_beamform (cb_tr_k, then the "
                       "tbtf_k reorder)"),
            "window_achieved": round(window_tops, 2),
        },
    }

    gpu_weights = d_w.download()      # the last block's weights, for the CPU baseline's check
    # the epoch of the last weights written to d_w (the fused launches above
    # when the block uses them, else the per-kernel pass)
    gpu_weights_epoch = w_epoch[0]
    wp.close()
    cb.close()
    ib.close()
    for b in (*d_taftps, d_ftpa, d_delays, d_w, d_cb, d_ib):
        if b is not None:
            b.free()
    if not args.no_extra_legs:
        result["pipeline_h2d_inclusive"] = pipeline_leg()
        result["weights_generation"] = weights_leg(args.warmup)
        result["config1_block"] = config1_leg(args.warmup)
    if not args.no_mfma_reference:
        ref = mfma_reference_leg()
        if ref is not None:
            ref["cb_kernel_over_reference"] = round(cb_tops / ref["tops"], 4)
        result["roofline"]["practical_reference"] = ref
    if not args.no_corner_turns:
        result["corner_turns"] = corner_turn_leg(min(K, 100), args.warmup)
    if not args.no_batch:
        result["batched_subarrays"] = batch_leg(min(K, 100), args.warmup)
    if not args.no_serving:
        result["serving"] = serving_leg(args.serving_seconds, int(args.hbm_budget_gb * 1e9))
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cfg, sub, taftp, args.cpu_channels, delays, freqs, gpu_weights,
                                              gpu_weights_epoch)
    print(json.dumps(result), flush=True)
    return 0


PIPELINE_BENCH = os.path.join(ROOT, "psrdada_cpp_amd", "bin", "fbfuse_pipeline_bench")


def profile_leg(args):
    """One leg on its own (device 0), its JSON on one line."""
    from psrdada_cpp_amd import capi

    capi.call("fbf_set_device", 0)
    K = min(args.steps, 100)
    legs = {"batch": lambda: batch_leg(K, args.warmup),
            "corner_turns": lambda: corner_turn_leg(K, args.warmup),
            "serving": lambda: serving_leg(args.serving_seconds, int(args.hbm_budget_gb * 1e9)),
            "pipeline": pipeline_leg,
            "weights": lambda: weights_leg(args.warmup),
            "config1": lambda: config1_leg(args.warmup)}
    print(json.dumps({args.profile_leg: legs[args.profile_leg]()}), flush=True)
    return 0


def pipeline_leg(timed=40, warmup=8):
    """The drop-in C++ Pipeline at the same geometry with the voltages
    crossing PCIe every block (pinned host ring blocks, three streams:
    H2D, processing, D2H), run as a child process of its own (the C++
    layer's benchmark tool).  None if the tool is not built."""
    import subprocess

    if not os.path.exists(PIPELINE_BENCH):
        print("pipeline leg skipped: fbfuse_pipeline_bench not built", file=sys.stderr)
        return None
    try:
        r = subprocess.run([PIPELINE_BENCH, str(timed), str(warmup)], capture_output=True, text=True, timeout=300)
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        if r.returncode != 0 or not lines:
            print(f"pipeline leg failed: exit {r.returncode}: {r.stderr[-500:]}", file=sys.stderr)
            return None
        return json.loads(lines[-1])
    except (OSError, subprocess.SubprocessError, ValueError) as ex:
        print(f"pipeline leg skipped: {type(ex).__name__}: {ex}", file=sys.stderr)
        return None


def weights_leg(warmup, reps=50):
    """Delay -> weights for 64 antennas x 1024 beams x 64 channels (tau0 ~
    U(-2us, 2us), rate ~ U(-1e-9, 1e-9)), epochs spread over 600 s: HIP
    events around `reps` launches on one stream.  The accuracy against the
    FP64 oracle is checked in the CPU baseline."""
    from psrdada_cpp_amd import capi, workloads as wl

    A, B, F = 64, 1024, 64
    freqs = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F * wl.CHANNEL_BW_HZ, F)
    d = wl.synth_delays(B, A, 4242, 2e-6, rate_max=1e-9)
    s = capi.Stream()
    d_d = capi.DeviceBuffer.from_array(d, s)
    d_w = capi.DeviceBuffer(F * B * A * 2)
    wp = capi.WeightsPlan(A, B, F, freqs)
    for k in range(warmup):
        wp.generate(d_d, 600.0 * k / reps, d_w, stream=s)
    s.sync()
    e0, e1 = capi.Event(), capi.Event()
    e0.record(s)
    for k in range(reps):
        wp.generate(d_d, 600.0 * k / reps, d_w, stream=s)
    e1.record(s)
    s.sync()
    ms = e0.elapsed_ms(e1) / reps
    n = F * B * A
    block_ms = wl.config2().block_seconds * 1e3
    wp.close()
    d_d.free()
    d_w.free()
    return {"workload": "64 ant x 1024 beams x 64 ch, tau0 ~ U(-2us, 2us), rate ~ U(-1e-9, 1e-9), epochs over 600 s",
            "ms_per_launch": round(ms, 4), "weights_per_s": float(f"{n / (ms * 1e-3):.4g}"),
            "hbm_gbs": round((n * 2 + B * A * 8) / (ms * 1e-3) / 1e9, 1),
            "refresh_every_8_blocks_fraction_of_real_time": float(f"{ms / (8 * block_ms):.3g}")}


def config1_leg(warmup, reps=200):
    """The CPU plumbing case on the GPU: 16 antennas, 16 channels, 8 heap
    groups (2048 spectra = 9.8 ms), 32 beams, tscrunch 16: weights,
    split-transpose + incoherent beam, coherent beam per block, HIP events
    around `reps` blocks."""
    from psrdada_cpp_amd import capi, workloads as wl

    cfg = wl.config1()
    sub = cfg.subarrays[0]
    A, F, B = sub.nantennas, cfg.nchans, sub.nbeams
    (cb_off, cb_sc), (ib_off, ib_sc) = cfg.levels(sub)
    s = capi.Stream()
    d_t = capi.DeviceBuffer.from_array(wl.synth_taftp(cfg, 11), s)
    d_d = capi.DeviceBuffer.from_array(wl.synth_delays(B, A, 12, cfg.tau_max_s), s)
    d_f = capi.DeviceBuffer(cfg.ftpa_bytes(sub))
    d_w = capi.DeviceBuffer(cfg.weights_bytes(sub))
    wp = capi.WeightsPlan(A, B, F, cfg.channel_frequencies())
    cb = capi.CbPlan(A, B, F, sub.tscrunch, sub.fscrunch, cb_off, cb_sc, nsamples_per_heap_out=cfg.cb_ohs(sub))
    ib = capi.IbPlan(A, F, cfg.ib_tscrunch, cfg.ib_fscrunch, ib_off, ib_sc)
    d_cb = capi.DeviceBuffer(cfg.cb_out_bytes(sub))
    d_ib = capi.DeviceBuffer(cfg.ib_out_bytes())

    def block(k):
        wp.generate(d_d, cfg.block_seconds * k, d_w, stream=s)
        capi.split_transpose_ib(d_t, d_f, F, ib, d_ib, stream=s)
        cb.beamform(d_f, d_w, d_cb, stream=s)

    for k in range(warmup):
        block(k)
    s.sync()
    e0, e1 = capi.Event(), capi.Event()
    e0.record(s)
    for k in range(reps):
        block(k)
    e1.record(s)
    s.sync()
    ms = e0.elapsed_ms(e1) / reps
    for o in (wp, cb, ib):
        o.close()
    return {"workload": cfg.name, "device_ms_per_block": round(ms, 4),
            "x_real_time": round(cfg.block_seconds * 1e3 / ms, 1), "block_ms": round(cfg.block_seconds * 1e3, 3)}


TRAFFIC_PROFILES = {
    "ftpa": os.path.join(ROOT, "profiles", "cb_64ant_64ch_512beams_hbm_traffic.txt"),
    "fused": os.path.join(ROOT, "profiles", "cb_ib_fused_64ant_64ch_512beams_hbm_traffic.txt"),
}


def profiled_traffic(path):
    """HBM bytes per launch of cb_tr_k on this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summary committed under profiles/ (counter
    readings scaled by what they read for a plain 1 GiB copy); None when
    there is no summary for the path."""
    if path not in TRAFFIC_PROFILES:
        return None
    try:
        with open(TRAFFIC_PROFILES[path]) as f:
            for line in f:
                if line.startswith("kernel_traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def calibrate_copy():
    """A plain elementwise copy (int32 negate) of exactly 1 GiB in and 1 GiB
    out, launched 3 times: what the HBM byte counters read for it converts
    their units into bytes."""
    import torch

    n = 1 << 28
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy": {"kernel": "torch.neg int32", "read_bytes": 4 * n, "write_bytes": 4 * n,
                                         "launches": 3}}), flush=True)
    return 0


def mfma_reference_leg():
    """What the vendor's int8 GEMM (hipBLASLt through torch._int_mm) reaches
    on this GPU in the same run, for uniformly random int8 operands
    (16384 x 16384 x 8192, int32 out): the int8 MFMA rate depends on the
    operand data (zeros run about 25 % faster), so on beamformer-like data
    the nominal 5.0 POP/s is not reachable and this is the practical
    reference beside it.  It runs in a child process: torch carries its own
    HIP runtime, which finds no device in a process whose other HIP runtime
    (the beamformer library's) already holds it.  None if unavailable."""
    import subprocess

    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--mfma-reference-only"],
                           capture_output=True, text=True, timeout=300)
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        if r.returncode != 0 or not lines:
            print(f"mfma reference leg skipped: exit {r.returncode}: {r.stderr[-500:]}", file=sys.stderr)
            return None
        return json.loads(lines[-1])
    except (OSError, subprocess.SubprocessError, ValueError) as ex:
        print(f"mfma reference leg skipped: {type(ex).__name__}: {ex}", file=sys.stderr)
        return None


def mfma_reference():
    """The vendor int8 GEMM measurement of mfma_reference_leg (in this
    process), or None."""
    try:
        import torch

        m, n, k = 16384, 16384, 8192
        g = torch.Generator(device="cuda").manual_seed(5)
        a = torch.randint(-128, 128, (m, k), dtype=torch.int8, device="cuda", generator=g)
        b = torch.randint(-128, 128, (k, n), dtype=torch.int8, device="cuda", generator=g)
        for _ in range(3):
            c = torch._int_mm(a, b)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record()
        for _ in range(reps):
            c = torch._int_mm(a, b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del a, b, c
        torch.cuda.empty_cache()
        return {"what": "hipBLASLt int8 GEMM (torch._int_mm), 16384 x 16384 x 8192, uniform random int8, same run",
                "kernel_ms": round(ms, 4), "tops": round(2.0 * m * n * k / (ms * 1e-3) / 1e12, 1)}
    except Exception as ex:  # noqa: BLE001 - a missing reference is reported, not fatal
        print(f"mfma reference leg skipped: {type(ex).__name__}: {ex}", file=sys.stderr)
        return None


def corner_turn_leg(K, warmup):
    """The byte reorders either side of the beamformer, device-resident, HIP
    events around K launches on one stream: the TUSE beam corner turn of 512
    beams x 16 bands x 512 samples (4 groups of 128) x 64 channels (256 MiB),
    and the F-engine heap reorder of 1024 timestamps x 256 channels x 256
    samples x 2 pols (256 MiB).  HBM-bound: each byte read once and written
    once, so the algorithmic traffic is twice the block."""
    from psrdada_cpp_amd import capi, workloads as wl

    out = {}
    s = capi.Stream()
    C, S, F, B, G = 64, 128, 16, 512, 4
    n = G * B * F * S * C
    d_in = capi.DeviceBuffer.from_array(capi.synth_gaussian_int8(7, 30.0, n), s)
    d_out = capi.DeviceBuffer(n)
    # reference: a device-to-device copy of the same bytes (the HIP runtime's
    # copy kernel), the practical ceiling of a read-once write-once kernel
    legs = [("device_copy", lambda: capi.call("fbf_memcpy_d2d", d_out.ptr, d_in.ptr, n, s.handle),
             "hipMemcpyAsync device to device of the same 256 MiB (ceiling reference)"),
            ("tuse_corner_turn", lambda: capi.tuse_transpose(d_in, d_out, C, S, F, B, G, stream=s),
             "512 beams x 16 bands x 512 samples (4 groups of 128) x 64 channels, per-beam [group][sample][band][chan]"),
            ("feng_to_dada", lambda: capi.feng_to_dada(d_in, d_out, 256, stream=s),
             "1024 timestamps of [256 channels][256 samples] x 2 pols char2 -> [256 samples][256 channels]")]
    for name, fn, workload in legs:
        for _ in range(warmup):
            fn()
        s.sync()
        e0, e1 = capi.Event(), capi.Event()
        e0.record(s)
        for _ in range(K):
            fn()
        e1.record(s)
        s.sync()
        ms = e0.elapsed_ms(e1) / K
        gbs = 2 * n / (ms * 1e-3) / 1e9
        out[name] = {"workload": workload, "bytes_per_launch": 2 * n, "kernel_ms": round(ms, 4),
                     "achieved_gbs": round(gbs, 1), "peak_gbs": wl.PEAK_HBM_GBS,
                     "frac": round(gbs / wl.PEAK_HBM_GBS, 4)}
    d_in.free()
    d_out.free()
    return out


def batch_leg(K, warmup):
    """One 64-antenna block split 40/16/8 antennas with 768/256/64 beams, plus
    the incoherent beam, as one captured hipGraph per block; kernel variants
    and chunk lengths picked by the autotuner, for tscrunch 4, 16 and 64."""
    from psrdada_cpp_amd import capi, workloads as wl

    out = {}
    for ts in (4, 16, 64):
        cfg = wl.config3(ts)
        subs = [dict(start_antenna=sb.start_antenna, nantennas=sb.nantennas, nbeams=sb.nbeams, tscrunch=ts,
                     fscrunch=1, nsamples_per_heap_out=cfg.nsamples // ts) for sb in cfg.subarrays]
        levels = [cfg.levels(sb)[0] for sb in cfg.subarrays]
        batch = capi.Batch(cfg.total_nantennas, cfg.nchans, cfg.nheap_groups, subs, levels)
        (_, _), (ibo, ibs) = cfg.levels(cfg.subarrays[0])
        batch.enable_ib(cfg.ib_tscrunch, cfg.ib_fscrunch, ibo, ibs)
        t_tune = time.perf_counter()
        batch.autotune(reps=3)
        t_tune = time.perf_counter() - t_tune
        s = capi.Stream()
        d_taftp = capi.DeviceBuffer.from_array(wl.synth_taftp(cfg, 31 + ts), s)
        ws = []
        for i, sb in enumerate(cfg.subarrays):
            wpi = capi.WeightsPlan(sb.nantennas, sb.nbeams, cfg.nchans, cfg.channel_frequencies())
            d_wi = capi.DeviceBuffer(batch.weights_bytes(i))
            wpi.generate(capi.DeviceBuffer.from_array(wl.synth_delays(sb.nbeams, sb.nantennas, 40 + i,
                                                                      cfg.tau_max_s), s), 0.0, d_wi, stream=s)
            s.sync()
            wpi.close()
            ws.append(d_wi)
        outs = [capi.DeviceBuffer(batch.output_bytes(i)) for i in range(len(subs) + 1)]
        for _ in range(warmup):
            batch.run(d_taftp, ws, outs, use_graph=True, stream=s)
        s.sync()
        e0, e1 = capi.Event(), capi.Event()
        t0 = time.perf_counter()
        e0.record(s)
        for _ in range(K):
            batch.run(d_taftp, ws, outs, use_graph=True, stream=s)
        e1.record(s)
        s.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3 / K
        dev_ms = e0.elapsed_ms(e1) / K
        ops = cfg.cb_ops()
        out[f"tscrunch_{ts}"] = {
            "ms_per_block": round(wall_ms, 4),
            "device_ms_per_block": round(dev_ms, 4),
            "x_real_time": round(cfg.block_seconds * 1e3 / wall_ms, 2),
            "int8_tops": round(ops / (wall_ms * 1e-3) / 1e12, 2),
            "autotune_s": round(t_tune, 2),
            "schedule": "in turn on one stream" if batch.serial else "subarrays side by side",
        }
        batch.close()
        for b in (d_taftp, *ws, *outs):
            b.free()
    out["workload"] = ("64-ant block split 40/16/8 ant with 768/256/64 beams + IB, one hipGraph per block; "
                       "the autotuner picks each subarray's kernel and the block's schedule")
    return out


def serving_leg(seconds, budget):
    """Four feeds (64/768/64, 32/256/64, 16/128/32, 8/64/16 ant/beams/ch),
    Poisson arrivals at 1.5x real time per feed from pinned host memory, then
    closed-loop saturation with host and with device-resident input for 2 s
    each; and the 64 ant/512 beams/64 ch geometry as a feed of its own,
    device-resident, closed loop for 2.5 s (throughput and p99 through the
    engine).  One untimed warm-up block per feed precedes each run."""
    from psrdada_cpp_amd import serving

    feeds = serving.four_subarray_feeds(32)
    srv, fids = serving.make_server(feeds, budget, nslots=8)
    try:
        out = {
            "workload": "4 feeds 64/768/64, 32/256/64, 16/128/32, 8/64/16 (ant/beams/ch), 32 heap groups, "
                        "ts 16 + IB; continuous batching over 8 streams",
            "poisson_1.5x": serving.run_poisson(srv, fids, feeds, 1.5, seconds),
            "saturation_host_input": serving.run_saturation(srv, fids, feeds, seconds=2.0),
            "saturation_device_input": serving.run_saturation(srv, fids, feeds, device_input=True, seconds=2.0),
        }
    finally:
        srv.close()
    out["block_ms"] = round(feeds[0].block_seconds * 1e3, 3)
    one = [serving.FeedSpec(64, 512, 64, 32)]
    srv, fids = serving.make_server(one, budget, nslots=8)
    try:
        r = serving.run_saturation(srv, fids, one, device_input=True, seconds=2.5, nout=8)
        r["workload"] = "one feed 64 ant/512 beams/64 ch, 32 heap groups, ts 16 + IB, device-resident input"
        out["feed_64ant_512beams_64ch_device_input"] = r
    finally:
        srv.close()
    return out


def cpu_baseline(cfg, sub, taftp, nch, delays, freqs, gpu_weights, gpu_weights_epoch):
    """The CPU oracle on a bounded sample of the same block: split-transpose
    and IB over the whole block, CB (exact power + quantiser) over `nch` of the
    64 channels, scaled linearly in F to a full block.  Also the CPU side of
    config 1 (a whole block) and config 4 (the FP64 weights generator), and,
    as the checker, the GPU weights of the last benchmarked block and of the
    config-4 workload against the FP64 oracle."""
    import oracle
    from psrdada_cpp_amd import capi, workloads as wl

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    A, F, B, T = sub.nantennas, cfg.nchans, sub.nbeams, cfg.nsamples
    nhg = cfg.nheap_groups
    t0 = time.perf_counter()
    ftpa = oracle.split_transpose(taftp, nhg, cfg.total_nantennas, A, 0, F)
    t1 = time.perf_counter()
    pib = oracle.ib_power(taftp, F, cfg.ib_tscrunch, cfg.ib_fscrunch, nhg, cfg.total_nantennas, nthreads=cores)
    t2 = time.perf_counter()
    w = capi.synth_gaussian_int8(7, 60.0, nch * B * A * 2)
    v = ftpa[: nch * T * 2 * A * 2]
    p = oracle.cb_power(v, w, nch, T, A, B, sub.tscrunch, sub.fscrunch, nthreads=cores)
    q = oracle.cb_tbtf(p, np.ones(nch, np.float32), np.ones(nch, np.float32), T // sub.tscrunch)
    t3 = time.perf_counter()
    del q
    block_cpu_s = (t1 - t0) + (t2 - t1) + (t3 - t2) * (F / nch)
    # One core: IB over the whole block and CB over 4 channels, 1 thread.
    n1 = min(4, nch)
    u0 = time.perf_counter()
    oracle.ib_power(taftp, F, cfg.ib_tscrunch, cfg.ib_fscrunch, nhg, cfg.total_nantennas, nthreads=1)
    u1 = time.perf_counter()
    p1 = oracle.cb_power(ftpa[: n1 * T * 2 * A * 2], w[: n1 * B * A * 2], n1, T, A, B, sub.tscrunch, sub.fscrunch,
                         nthreads=1)
    oracle.cb_tbtf(p1, np.ones(n1, np.float32), np.ones(n1, np.float32), T // sub.tscrunch)
    u2 = time.perf_counter()
    block_1core_s = (t1 - t0) + (u1 - u0) + (u2 - u1) * (F / n1)
    # Config 1: a whole block on the CPU (split-transpose, IB, CB + quantiser).
    c1 = wl.config1()
    s1 = c1.subarrays[0]
    t1_in = wl.synth_taftp(c1, 11)
    w1 = oracle.weights_f64(wl.synth_delays(s1.nbeams, s1.nantennas, 12, c1.tau_max_s), c1.channel_frequencies(),
                            0.0, s1.nantennas, s1.nbeams, c1.nchans)
    (o1, c1s), (io1, is1) = c1.levels(s1)
    v0 = time.perf_counter()
    f1 = oracle.split_transpose(t1_in, c1.nheap_groups, s1.nantennas, s1.nantennas, 0, c1.nchans)
    oracle.ib_tf(oracle.ib_power(t1_in, c1.nchans, c1.ib_tscrunch, 1, c1.nheap_groups, s1.nantennas, nthreads=cores),
                 io1, is1)
    oracle.cb_tbtf(oracle.cb_power(f1, w1, c1.nchans, c1.nsamples, s1.nantennas, s1.nbeams, s1.tscrunch, 1,
                                   nthreads=cores), o1, c1s, c1.cb_ohs(s1))
    c1_ms = (time.perf_counter() - v0) * 1e3
    # Config 4: the FP64 weights generator on the CPU (one thread), and the
    # GPU's weights against it.
    A4, B4, F4 = 64, 1024, 64
    fr4 = capi.channel_frequencies_f64(wl.L_BAND_CENTRE_HZ, F4 * wl.CHANNEL_BW_HZ, F4)
    d4 = wl.synth_delays(B4, A4, 4242, 2e-6, rate_max=1e-9)
    x0 = time.perf_counter()
    w4_ref, frac4 = oracle.weights_f64(d4, fr4, 300.0, A4, B4, F4, want_phase=True)
    w4_s = time.perf_counter() - x0
    w4_gpu, ph4_gpu = _gpu_weights(d4, fr4, 300.0, A4, B4, F4)
    dw4 = np.abs(w4_gpu.astype(int) - w4_ref.astype(int))
    wref2 = oracle.weights_f64(delays, freqs, gpu_weights_epoch, sub.nantennas, sub.nbeams, F)
    dw2 = np.abs(gpu_weights.astype(int) - wref2.astype(int))
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": round(cfg.block_seconds / block_cpu_s, 4),
        "unit": "x real-time",
        "cores": cores,
        "kind": "port",
        "sample": f"one config-2 block: split-transpose (1 thread) and IB over all 64 channels, "
                  f"CB over {nch} of 64 channels scaled x{F // nch}; "
                  f"measured {t3 - t0:.1f} s, extrapolated block time {block_cpu_s:.2f} s",
        "int8_tops": round(cfg.cb_ops() / block_cpu_s / 1e12, 4),
        "cb_seconds_per_block": round((t3 - t2) * (F / nch), 3),
        "cpu_model": cpu_model,
        "host_cpus": os.cpu_count(),
        "one_core": {"value": round(cfg.block_seconds / block_1core_s, 5), "unit": "x real-time",
                     "sample": f"split-transpose and IB over the whole block, CB over {n1} channels scaled "
                               f"x{F // n1}, 1 thread; extrapolated block time {block_1core_s:.1f} s"},
        "config1_block": {"ms_per_block": round(c1_ms, 3), "x_real_time": round(c1.block_seconds * 1e3 / c1_ms, 3),
                          "threads": cores, "sample": "one whole config-1 block (16 ant, 16 ch, 2048 spectra, "
                                                      "32 beams): split-transpose, IB, CB and quantiser"},
        "config4_weights": {"cpu_fp64_seconds": round(w4_s, 3), "cpu_weights_per_s": float(f"{A4 * B4 * F4 / w4_s:.4g}"),
                            "threads": 1,
                            "gpu_vs_fp64_oracle": {
                                "max_phase_error_rad": float(f"{np.abs(ph4_gpu.astype(np.float64) - frac4).max() * 2 * np.pi:.3g}"),
                                "fraction_1_lsb": float(f"{(dw4 > 0).mean():.3g}"), "max_lsb": int(dw4.max()),
                                "epoch_s": 300.0},
                            "config2_last_block_fraction_1_lsb": float(f"{(dw2 > 0).mean():.3g}"),
                            "config2_last_block_max_lsb": int(dw2.max())},
    }


def _gpu_weights(delays, freqs, epoch, A, B, F):
    """GPU weights and reduced phases for the CPU baseline's check."""
    from psrdada_cpp_amd import capi

    wp = capi.WeightsPlan(A, B, F, freqs)
    d_d = capi.DeviceBuffer.from_array(delays)
    d_w = capi.DeviceBuffer(F * B * A * 2)
    d_ph = capi.DeviceBuffer(F * B * A * 4)
    wp.generate(d_d, epoch, d_w, d_ph)
    out = d_w.download(), d_ph.download(np.float32)
    wp.close()
    for b in (d_d, d_w, d_ph):
        b.free()
    return out


if __name__ == "__main__":
    sys.exit(main())
