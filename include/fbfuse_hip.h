/*
 * fbfuse_hip.h -- C ABI of the MI355X-native FBFUSE beamformer.
 *
 * Plain C: pointers, sizes and status codes; no HIP, thrust or torch types in
 * any signature.  Streams are passed as opaque `fbf_stream_t` (a hipStream_t;
 * NULL is the default stream).  Every function returns FBF_OK (0) or a
 * non-zero status; fbf_last_error() gives the message of the last failure on
 * the calling thread.  No exception crosses the ABI.
 *
 * Compute entry points are ASYNCHRONOUS on the given stream (they only
 * enqueue work); the C++ operator classes above this ABI synchronise the
 * stream on return, which restores the reference's synchronous contract.
 *
 * Data layouts (char2 = one complex int8 sample, re then im; npol = 2):
 *   TAFTP voltages  [heap_group][antenna][channel][sample_in_heap][pol] char2
 *   FTPA voltages   [channel][sample][pol][antenna] char2
 *   FBA weights     [channel][beam][antenna] char2
 *   TBTF CB output  [sample_out / ohs][beam][sample_out % ohs][channel_out] int8
 *   TF IB output    [sample_out][channel_out] int8
 *   delays          [beam][antenna] float2 (x = rate in s/s, y = offset in s)
 *
 * Reference interfaces replaced (paths relative to the psrdada_cpp tree):
 *   fbf_split_transpose   <- SplitTranspose::transpose, SplitTranspose.cuh:73-74,
 *                            kernel SplitTranspose.cu:11-70
 *   fbf_cb_*              <- CoherentBeamformer::beamform, CoherentBeamformer.cuh:128-131,
 *                            kernel CoherentBeamformer.cu:48-170
 *   fbf_ib_*              <- IncoherentBeamformer::beamform, IncoherentBeamformer.cuh:52-54,
 *                            kernel IncoherentBeamformer.cu:10-69
 *   fbf_weights_*         <- WeightsManager::weights, WeightsManager.cuh:69,
 *                            kernel WeightsManager.cu:14-79
 *   fbf_power_levels      <- PipelineConfig::update_power_offsets_and_scalings,
 *                            PipelineConfig.cpp:109-127
 *   fbf_channel_frequencies <- PipelineConfig::calculate_channel_frequencies,
 *                            PipelineConfig.cpp:180-195
 *   fbf_malloc/fbf_memcpy_* <- thrust::device_vector storage used by
 *                            DoubleDeviceBuffer, double_device_buffer.cuh:9-10
 */
#ifndef FBFUSE_HIP_H
#define FBFUSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* fbf_stream_t;
typedef void* fbf_event_t;

enum fbf_status {
    FBF_OK = 0,
    FBF_ERR_ARG = 1,          /* bad argument (null pointer, bad enum, bad shape) */
    FBF_ERR_SIZE = 2,         /* buffer size does not match the geometry */
    FBF_ERR_HIP = 3,          /* HIP runtime error */
    FBF_ERR_UNSUPPORTED = 4,  /* geometry outside what the kernels support */
    FBF_ERR_STATE = 5,        /* call not valid in the object's current state */
    FBF_ERR_CAPACITY = 6      /* admission refused: budget or queue full */
};

/* Quantiser of the detected power (CoherentBeamformer.cu:169):
 * ROUND: clamp(round_half_away((P - offset) / scale), -128, 127)
 * TRUNC: clamp(trunc((P - offset) / scale), -128, 127) -- the reference cast.
 * (P - offset) / scale is evaluated in IEEE double from the exact power. */
enum fbf_quant_mode { FBF_QUANT_ROUND = 0, FBF_QUANT_TRUNC = 1 };

const char* fbf_last_error(void);
int fbf_version(void);

/* ---- device plumbing (hipMalloc & co; for callers without their own HIP) */
int fbf_device_count(int* count);
int fbf_set_device(int device);
int fbf_device_name(char* buf, size_t buflen);
int fbf_malloc(void** ptr, size_t bytes);
int fbf_free(void* ptr);
int fbf_host_malloc(void** ptr, size_t bytes);      /* pinned host memory */
int fbf_host_free(void* ptr);
int fbf_memcpy_h2d(void* dst, const void* src, size_t bytes, fbf_stream_t stream);
int fbf_memcpy_d2h(void* dst, const void* src, size_t bytes, fbf_stream_t stream);
int fbf_memcpy_d2d(void* dst, const void* src, size_t bytes, fbf_stream_t stream);
int fbf_memset(void* dst, int value, size_t bytes, fbf_stream_t stream);
int fbf_stream_create(fbf_stream_t* stream);
int fbf_stream_destroy(fbf_stream_t stream);
int fbf_stream_synchronize(fbf_stream_t stream);
int fbf_device_synchronize(void);
int fbf_event_create(fbf_event_t* event);
int fbf_event_destroy(fbf_event_t event);
int fbf_event_record(fbf_event_t event, fbf_stream_t stream);
/* Work enqueued on `stream` after this call waits for `event` (its last
 * record) to complete: the cross-stream ordering of a pipeline that runs
 * the weights generator beside the split-transpose. */
int fbf_stream_wait_event(fbf_stream_t stream, fbf_event_t event);
int fbf_event_synchronize(fbf_event_t event);
int fbf_event_elapsed_ms(fbf_event_t start, fbf_event_t stop, float* ms);

/* ---- deterministic synthetic data (host, splitmix64 + Box-Muller) ------ */
/* re/im = round(N(0, sigma)); clip != 0: clamp to [-127, 127], else wrap to
 * int8 like static_cast<int8_t>(lround(x)).  Writes n bytes. */
int fbf_synth_gaussian_int8(uint64_t seed, double sigma, int clip, int8_t* out, size_t n);
/* n floats uniform in [lo, hi). */
int fbf_synth_uniform_f32(uint64_t seed, float lo, float hi, float* out, size_t n);

/* ---- configuration helpers -------------------------------------------- */
/* PipelineConfig.cpp:109-127 (same float arithmetic). */
int fbf_power_levels(float input_level, float output_level,
                     int cb_nantennas, int cb_tscrunch, int cb_fscrunch,
                     int ib_nantennas, int ib_tscrunch, int ib_fscrunch, int npol,
                     float* cb_offset, float* cb_scale,
                     float* ib_offset, float* ib_scale);
/* PipelineConfig.cpp:180-195 (float) and the same centres in double. */
int fbf_channel_frequencies(float cfreq, float bw, int nchans, float* freqs);
int fbf_channel_frequencies_f64(double cfreq, double bw, int nchans, double* freqs);

/* ---- split transpose (SplitTranspose.cu:11-126) ------------------------ */
/* TAFTP -> FTPA for the contiguous antenna subset
 * [start_antenna, start_antenna + used_nantennas). ftpa_bytes must equal
 * 2*2*used_nantennas*nchans*nheap_groups*nsamples_per_heap. */
int fbf_split_transpose(const void* taftp, size_t taftp_bytes,
                        void* ftpa, size_t ftpa_bytes,
                        int total_nantennas, int used_nantennas, int start_antenna,
                        int nchans, int nsamples_per_heap, fbf_stream_t stream);

/* ---- coherent beamformer (CoherentBeamformer.cu:48-225) ---------------- */
typedef struct fbf_cb_desc {
    int nantennas;              /* antennas in the FTPA input (A) */
    int nbeams;                 /* beams (B), multiple of 8 */
    int nchans;                 /* channels (F) */
    int tscrunch;               /* time integration */
    int fscrunch;               /* channel integration */
    int nsamples_per_heap_out;  /* ohs: output samples per TBTF heap; 0 -> 8192 / (F / fscrunch)
                                   (FBFUSE_CB_NSAMPLES_PER_HEAP, fbfuse_constants.hpp:19-25) */
    int quant_mode;             /* enum fbf_quant_mode */
} fbf_cb_desc;

typedef struct fbf_cb_plan fbf_cb_plan;

/* offsets/scales: host arrays of F / fscrunch floats (per output channel). */
int fbf_cb_plan_create(const fbf_cb_desc* desc, const float* offsets, const float* scales,
                       fbf_cb_plan** plan);
/* New per-channel levels for later calls.  Waits for all work on the device
 * first (a beamform in flight finishes with the old levels), so it must not
 * be called while a stream is being captured. */
int fbf_cb_plan_set_levels(fbf_cb_plan* plan, const float* offsets, const float* scales);
int fbf_cb_plan_destroy(fbf_cb_plan* plan);
/* Kernel tile variant (0 = autotuned default); see fbf_cb_variant_count. */
int fbf_cb_plan_set_variant(fbf_cb_plan* plan, int variant);
int fbf_cb_variant_count(int* count);
int fbf_cb_output_bytes(const fbf_cb_plan* plan, size_t ftpa_bytes, size_t* tbtf_bytes);
/* FTPA + FBA -> int8 TBTF.  The int8 beams pass through a staging buffer
 * ([Fo][To][B], the size of the output) that the plan owns and grows on
 * demand, so calls with one plan on different streams must not overlap, and
 * a stream that is being captured needs fbf_cb_plan_reserve first. */
int fbf_cb_beamform(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes,
                    const void* fba_weights, size_t weights_bytes,
                    void* tbtf, size_t tbtf_bytes, fbf_stream_t stream);
/* As fbf_cb_beamform with a caller-owned staging buffer of at least
 * fbf_cb_output_bytes bytes (16-byte aligned): one plan can then run on
 * several streams at once.  No counterpart in the reference; used by the
 * serving engine. */
int fbf_cb_beamform_ws(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes,
                       const void* fba_weights, size_t weights_bytes,
                       void* tbtf, size_t tbtf_bytes, void* workspace, size_t workspace_bytes,
                       fbf_stream_t stream);
/* As fbf_cb_beamform, allocating nothing: the int8 beams are staged in a
 * fixed device array of the library (32 MiB per GPU, held until the process
 * ends) instead of a buffer the plan allocates.  No counterpart in the
 * reference; for callers that cannot allocate at call time (code under
 * stream capture, tight serving loops).  A staged output (the size of the
 * TBTF output) above 32 MiB returns FBF_ERR_SIZE.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int fbf_cb_beamform_noalloc(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes,
                            const void* fba_weights, size_t weights_bytes,
                            void* tbtf, size_t tbtf_bytes, fbf_stream_t stream);
/* Sizes the plan's staging buffer and the incoherent-beam partials of
 * fbf_cb_ib_beamform_taftp for blocks of nsamples samples (call before
 * capturing a stream that runs the plan).  The FTPA copy that
 * fbf_cb_beamform_taftp makes when it cannot stage the TAFTP block itself is
 * sized by its first call outside capture. */
int fbf_cb_plan_reserve(fbf_cb_plan* plan, long long nsamples);
/* Beamform the plan's antennas as a subarray of a full-antenna FTPA block:
 * ftpa_block holds block_nantennas antennas per column and this subarray's
 * antennas start at antenna_offset.  Same output as splitting the subset
 * out first (SplitTranspose + fbf_cb_beamform) without the copy.  No
 * counterpart in the reference; used by the batched multi-subarray path. */
int fbf_cb_beamform_subarray(fbf_cb_plan* plan, const void* ftpa_block, size_t block_bytes,
                             int block_nantennas, int antenna_offset,
                             const void* fba_weights, size_t weights_bytes,
                             void* tbtf, size_t tbtf_bytes, fbf_stream_t stream);
/* Beamform straight from a TAFTP block: replaces the pair
 * SplitTranspose::transpose + CoherentBeamformer::beamform as the
 * reference's Pipeline calls them (Pipeline.cu:166-168), with the same
 * output as fbf_split_transpose of antennas [start_antenna, start_antenna +
 * plan nantennas) of the block's nantennas_total, then fbf_cb_beamform.
 * Where fbf_cb_taftp_direct says so, the kernel stages the block's
 * [antenna][sample, pol] runs itself and transposes them in LDS, so no FTPA
 * copy is written; otherwise the call split-transposes into an FTPA buffer
 * that the plan owns (grown on demand, as the staging buffer) first.  As
 * with fbf_cb_beamform, calls with one plan on different streams must not
 * overlap (they share the plan's buffers).  With
 * the automatic variant and more than 48 antennas it picks the transposed
 * kernel with 128 beams per wave where that applies (half the transposed
 * LDS reads per MFMA of 64 beams per wave). */
int fbf_cb_beamform_taftp(fbf_cb_plan* plan, const void* taftp, size_t taftp_bytes,
                          int nantennas_total, int start_antenna, int nsamples_per_heap,
                          const void* fba_weights, size_t weights_bytes,
                          void* tbtf, size_t tbtf_bytes, fbf_stream_t stream);
/* *direct = 1 when fbf_cb_beamform_taftp reads a block of this geometry
 * without the FTPA copy: the plan's kernel is the transposed one with 64 or
 * 128 beams per wave, fscrunch 1, nsamples_per_heap a power of two >= 16,
 * and one heap group under 2 GiB.  No counterpart in the reference. */
int fbf_cb_taftp_direct(const fbf_cb_plan* plan, int nantennas_total, int nsamples_per_heap,
                        long long nsamples, int* direct);
/* Timing of the coherent kernel alone: every later beamform call with this
 * plan records `before` on its stream right before the coherent kernel's
 * launch and `after` right after it (before the TBTF reorder and anything
 * in that launch).  NULL, NULL stops it.  No counterpart in the reference;
 * for benchmarks and profilers. */
int fbf_cb_plan_set_timing(fbf_cb_plan* plan, fbf_event_t before, fbf_event_t after);
/* LDS chunk length (columns of the integration window staged per step),
 * 0 = automatic; a power of two <= 16 dividing 2*tscrunch. */
int fbf_cb_plan_set_chunk(fbf_cb_plan* plan, int chunk);
/* Times every kernel variant and chunk length on synthetic data of the
 * plan's geometry and nsamples samples, keeps the fastest on the plan and
 * in a process-wide table (later plans of the same geometry start from it).
 * Needs scratch device memory for one block.  best_* may be NULL. */
int fbf_cb_autotune(fbf_cb_plan* plan, long long nsamples, int reps,
                    int* best_variant, int* best_chunk, float* best_ms);
/* Exact int32 beam sums, int32 [F][T][P][B][2] (re, im). */
int fbf_cb_beam_sums(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes,
                     const void* fba_weights, size_t weights_bytes,
                     int32_t* sums, size_t sums_bytes, fbf_stream_t stream);
/* Exact detected power before quantisation, double [F/fs][T/ts][B]. */
int fbf_cb_beam_power(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes,
                      const void* fba_weights, size_t weights_bytes,
                      double* power, size_t power_bytes, fbf_stream_t stream);

/* ---- incoherent beamformer (IncoherentBeamformer.cu:10-117) ------------ */
typedef struct fbf_ib_desc {
    int nantennas;              /* all antennas of the TAFTP input */
    int nchans;
    int nsamples_per_heap;      /* 256 */
    int tscrunch;               /* divides nsamples_per_heap */
    int fscrunch;
    int quant_mode;
} fbf_ib_desc;

typedef struct fbf_ib_plan fbf_ib_plan;

int fbf_ib_plan_create(const fbf_ib_desc* desc, const float* offsets, const float* scales,
                       fbf_ib_plan** plan);
/* As fbf_cb_plan_set_levels: waits for all work on the device first. */
int fbf_ib_plan_set_levels(fbf_ib_plan* plan, const float* offsets, const float* scales);
int fbf_ib_plan_destroy(fbf_ib_plan* plan);
int fbf_ib_output_bytes(const fbf_ib_plan* plan, size_t taftp_bytes, size_t* tf_bytes);
int fbf_ib_beamform(fbf_ib_plan* plan, const void* taftp, size_t taftp_bytes,
                    void* tf, size_t tf_bytes, fbf_stream_t stream);
/* Exact power, double [T/ts][F/fs]. */
int fbf_ib_power(fbf_ib_plan* plan, const void* taftp, size_t taftp_bytes,
                 double* power, size_t power_bytes, fbf_stream_t stream);
/* Coherent and incoherent beams of one TAFTP block (all the coherent plan's
 * antennas; no subset) in one pass over it: outputs equal those of
 * fbf_cb_beamform_taftp(plan, ..., nantennas_total = plan nantennas,
 * start_antenna = 0, ...) and fbf_ib_beamform(ib_plan, ...).  Replaces the
 * reference Pipeline's three operator calls (Pipeline.cu:166-170).  Where
 * fbf_cb_ib_fused says so, the coherent kernel also leaves each sample's
 * power over the antennas and pols (from the operands it already holds)
 * in a buffer the coherent plan owns, and the first work-groups of the
 * TBTF reorder's launch scrunch and quantise them; otherwise the two
 * beamformers run one after the other.  Calls with one plan on different
 * streams must not overlap (they share the plan's buffers). */
int fbf_cb_ib_beamform_taftp(fbf_cb_plan* plan, fbf_ib_plan* ib_plan, const void* taftp, size_t taftp_bytes,
                             int nsamples_per_heap, const void* fba_weights, size_t weights_bytes,
                             void* tbtf, size_t tbtf_bytes, void* tf, size_t tf_bytes, fbf_stream_t stream);
/* *fused = 1 when fbf_cb_ib_beamform_taftp forms the incoherent beam inside
 * the coherent kernel: the direct TAFTP path with 128 beams per wave, and
 * an incoherent-beam plan of the same antennas, channels and heap size.  No
 * counterpart in the reference. */
int fbf_cb_ib_fused(const fbf_cb_plan* plan, const fbf_ib_plan* ib_plan, int nsamples_per_heap,
                    long long nsamples, int* fused);

/* Split-transpose of a full-antenna block (all antennas, no subset) and the
 * incoherent beam of the same block in one pass over the input: the FTPA
 * output equals fbf_split_transpose's and the TF output equals
 * fbf_ib_beamform's with ib_plan.  No counterpart in the reference (there the
 * two operators read the block separately); needs the plan's fscrunch to be
 * 1 and its tscrunch to divide 32, else FBF_ERR_UNSUPPORTED. */
int fbf_split_transpose_ib(const void* taftp, size_t taftp_bytes, void* ftpa, size_t ftpa_bytes,
                           int nchans, int nsamples_per_heap, fbf_ib_plan* ib_plan, void* tf,
                           size_t tf_bytes, fbf_stream_t stream);

/* ---- corner turns either side of the beamformer ------------------------ */

/* TUSE beam corner turn: replaces transpose::do_transpose
 * (meerkat/tuse/transpose_to_dada.hpp:20, meerkat/tuse/src/transpose_to_dada.cpp:19-48)
 * for the beams [beam0, beam0 + nbeams_out) at once.
 *   in  [ngroups][nbeams][nfreq][nsamples][nchans] bytes (in_bytes may exceed it)
 *   out [nbeams_out][ngroups][nsamples][nfreq][nchans] bytes: beam b's part,
 *       out + (b - beam0) * ngroups * nsamples * nfreq * nchans, is what
 *       do_transpose(..., beamnum = b, ...) writes.
 * Device pointers; asynchronous on `stream`. */
int fbf_tuse_transpose(const void* in, size_t in_bytes, void* out, size_t out_bytes, int nchans,
                       int nsamples, int nfreq, int nbeams, int ngroups, int beam0, int nbeams_out,
                       fbf_stream_t stream);

/* F-engine heap reorder: replaces kernels::feng_heaps_to_dada
 * (meerkat/tools/feng_to_dada.cuh:16-19, meerkat/tools/src/feng_to_dada.cu:54-80) as
 * launched by FengToDada::operator() (meerkat/tools/detail/feng_to_dada.cu:27-52).
 * Per timestamp, [nchans][256] 32-bit words (256 samples per heap x 2 pols x char2)
 * -> [256][nchans].  bytes must be a whole number of timestamps; any nchans > 0
 * (the reference kernel needs nchans % 32 == 0).  Out of place: in != out.
 * Device pointers; asynchronous on `stream`. */
int fbf_feng_to_dada(const void* in, void* out, size_t bytes, int nchans, fbf_stream_t stream);

/* ---- delay -> weights (WeightsManager.cu:14-123) ----------------------- */
typedef struct fbf_weights_plan fbf_weights_plan;

/* chan_freqs: host array of nchans channel centres in Hz (double). */
int fbf_weights_plan_create(int nantennas, int nbeams, int nchans, const double* chan_freqs,
                            fbf_weights_plan** plan);
int fbf_weights_plan_destroy(fbf_weights_plan* plan);
/* delays (device, [B][A] float2) at time t (s, FP64) -> weights [F][B][A] char2.
 * phase (optional device pointer, float [F][B][A]) receives the reduced
 * phase argument in cycles, phase - rint(phase), as passed to sin/cos. */
int fbf_weights_generate(fbf_weights_plan* plan, const void* delays, size_t delays_bytes,
                         double epoch, void* weights, size_t weights_bytes,
                         float* phase, fbf_stream_t stream);
/* Weights of ntsteps epochs tstart + k * tstep (k = 0 .. ntsteps - 1), written
 * [t][F][B][A]: generate_weights_k's tstart/tstep/ntsteps
 * (WeightsManager.cu:14-24, 58-73).  Step k equals fbf_weights_generate at
 * epoch tstart + k * tstep (FP64 product, then sum).  weights_bytes is
 * ntsteps * F * B * A * 2. */
int fbf_weights_generate_steps(fbf_weights_plan* plan, const void* delays, size_t delays_bytes,
                               double tstart, double tstep, int ntsteps, void* weights, size_t weights_bytes,
                               fbf_stream_t stream);
/* fbf_split_transpose_ib plus fbf_weights_generate(weights_plan, delays,
 * delays_bytes, epoch, weights, weights_bytes, NULL, stream) in one launch:
 * the launch's first work-groups generate the weights while the rest do the
 * transpose.  Outputs equal those of the two calls.  No counterpart in the
 * reference, whose Pipeline (Pipeline.cu:164-166) generates the weights and
 * then transposes in separate launches.  The weights plan's channels must
 * be the block's. */
int fbf_split_transpose_ib_weights(const void* taftp, size_t taftp_bytes, void* ftpa, size_t ftpa_bytes,
                                   int nchans, int nsamples_per_heap, fbf_ib_plan* ib_plan, void* tf,
                                   size_t tf_bytes, fbf_weights_plan* weights_plan, const void* delays,
                                   size_t delays_bytes, double epoch, void* weights, size_t weights_bytes,
                                   fbf_stream_t stream);
/* fbf_ib_beamform plus fbf_weights_generate(weights_plan, delays,
 * delays_bytes, epoch, weights, weights_bytes, NULL, stream) in one launch,
 * as fbf_split_transpose_ib_weights: the launch's first work-groups generate
 * the weights while the rest form the incoherent beam.  Outputs equal those
 * of the two calls.  It pairs with fbf_cb_beamform_taftp, which needs no
 * split-transpose pass; no counterpart in the reference (Pipeline.cu:164-170
 * launches the weights and the incoherent beam separately).  The weights
 * plan's channels must be the block's. */
int fbf_ib_beamform_weights(fbf_ib_plan* plan, const void* taftp, size_t taftp_bytes, void* tf,
                            size_t tf_bytes, fbf_weights_plan* weights_plan, const void* delays,
                            size_t delays_bytes, double epoch, void* weights, size_t weights_bytes,
                            fbf_stream_t stream);
/* fbf_cb_ib_beamform_taftp, and then fbf_weights_generate(weights_plan,
 * delays, delays_bytes, next_epoch, next_weights, next_weights_bytes, NULL,
 * stream): the next block's weights.  On the one-pass path the weights are
 * generated by the first work-groups of the TBTF reorder's launch, which
 * runs after the coherent kernel has read this block's weights, so
 * next_weights may be fba_weights itself.  Outputs equal those of the calls
 * in that order.  No counterpart in the reference, whose Pipeline generates
 * each block's weights at the start of that block (Pipeline.cu:164); a
 * caller reusing the next weights must regenerate them if the delay model
 * changes in between (the C++ Pipeline does). */
int fbf_cb_ib_beamform_taftp_next(fbf_cb_plan* plan, fbf_ib_plan* ib_plan, const void* taftp,
                                  size_t taftp_bytes, int nsamples_per_heap, const void* fba_weights,
                                  size_t weights_bytes, void* tbtf, size_t tbtf_bytes, void* tf, size_t tf_bytes,
                                  fbf_weights_plan* weights_plan, const void* delays, size_t delays_bytes,
                                  double next_epoch, void* next_weights, size_t next_weights_bytes,
                                  fbf_stream_t stream);

/* ---- batched subarrays on one block (no counterpart in the reference:
 * it compiles one subarray per process, fbfuse_constants.hpp:12-14) ---- */
typedef struct fbf_subarray_desc {
    int start_antenna;          /* first antenna of the contiguous subset */
    int nantennas;
    int nbeams;
    int tscrunch;
    int fscrunch;
    int nsamples_per_heap_out;  /* 0 -> 8192 / (F / fscrunch) */
    int quant_mode;
} fbf_subarray_desc;

typedef struct fbf_batch fbf_batch;

/* One TAFTP block of total_nantennas x nchans x nheap_groups heaps of
 * nsamples_per_heap samples, beamformed as nsub subarrays.  offsets[i] and
 * scales[i] are host arrays of nchans / subs[i].fscrunch floats. */
int fbf_batch_create(int total_nantennas, int nchans, int nsamples_per_heap, int nheap_groups,
                     int nsub, const fbf_subarray_desc* subs,
                     const float* const* offsets, const float* const* scales, fbf_batch** batch);
int fbf_batch_destroy(fbf_batch* batch);
/* Adds the incoherent beam over all antennas of the block (its output is
 * outputs[nsub] in fbf_batch_run). */
int fbf_batch_enable_ib(fbf_batch* batch, int tscrunch, int fscrunch, int quant_mode,
                        const float* offsets, const float* scales);
int fbf_batch_output_bytes(const fbf_batch* batch, int index, size_t* bytes);
int fbf_batch_weights_bytes(const fbf_batch* batch, int sub, size_t* bytes);
int fbf_batch_set_variant(fbf_batch* batch, int sub, int variant, int chunk);
/* Tunes each subarray's kernel (fbf_cb_autotune), then times whole blocks
 * with the subarrays in turn on the caller's stream (the default) and side
 * by side on their own streams, and switches to side by side only if it is
 * more than 3 % faster. */
int fbf_batch_autotune(fbf_batch* batch, int reps);
/* Schedule of a block: 1 = every kernel in turn on the caller's stream (the
 * default), 0 = subarrays side by side (fork/join over the batch's
 * streams).  No reference counterpart (the reference has one subarray per
 * pipeline). */
int fbf_batch_set_serial(fbf_batch* batch, int serial);
int fbf_batch_get_serial(const fbf_batch* batch, int* serial);
/* Split-transposes the block once into a full-antenna FTPA workspace and
 * runs every subarray's coherent beamformer on it (and the incoherent beam
 * on the TAFTP block), as independent branches forked from `stream` and
 * joined back into it.  use_graph != 0 captures this once into a hipGraph and
 * replays it; the graph is captured again whenever a buffer pointer
 * changes.  weights[i]: device FBA weights of subarray i; outputs: device
 * outputs (TBTF per subarray, then TF if the incoherent beam is enabled). */
int fbf_batch_run(fbf_batch* batch, const void* taftp, size_t taftp_bytes,
                  const void* const* weights, void* const* outputs, int use_graph, fbf_stream_t stream);

/* ---- serving engine (no counterpart in the reference, whose Pipeline
 * processes one subarray block by block, Pipeline.cu:173-247) -------------
 * Feeds are subarrays with their own geometry, plans and device-resident
 * weights.  Blocks submitted for any feed wait in one FIFO; a dispatcher
 * admits the oldest block as soon as a slot (a HIP stream) is free and its
 * buffers fit in the HBM budget, so blocks that arrive while others run are
 * admitted as capacity frees up and their work overlaps on the device.  Per
 * block: H2D copy (unless the input is already on the device), split
 * transpose, coherent beams, optional incoherent beam, D2H copy of the beams
 * into the caller's host buffers; completions are reported in finishing
 * order with their latency (submission to completion). */
typedef struct fbf_server fbf_server;

typedef struct fbf_feed_desc {
    int nantennas;              /* antennas of this feed's TAFTP blocks (all beamformed) */
    int nbeams;
    int nchans;
    int nheap_groups;           /* heap groups per block */
    int nsamples_per_heap;      /* 256 */
    int tscrunch;
    int fscrunch;
    int nsamples_per_heap_out;  /* 0 -> 8192 / (F / fscrunch) */
    int quant_mode;
    int ib_tscrunch;            /* 0: no incoherent beam */
    int ib_fscrunch;
} fbf_feed_desc;

typedef struct fbf_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t failed;
    uint64_t in_flight;
    uint64_t pending;
    double p50_ms;              /* latency percentiles over completed blocks */
    double p99_ms;
    double max_ms;
    double mean_ms;
    double input_bytes;         /* bytes of blocks completed */
    double wall_s;              /* seconds since fbf_server_start */
    size_t hbm_budget;
    size_t hbm_peak;            /* high-water mark of admitted block buffers */
} fbf_server_stats;

/* nslots concurrent blocks at most; hbm_budget bytes of device memory for
 * block buffers (allocated once, at fbf_server_start). */
int fbf_server_create(int nslots, size_t hbm_budget, fbf_server** server);
int fbf_server_destroy(fbf_server* server);
int fbf_server_add_feed(fbf_server* server, const fbf_feed_desc* desc,
                        const float* cb_offsets, const float* cb_scales,
                        const float* ib_offsets, const float* ib_scales, int* feed);
int fbf_server_feed_bytes(const fbf_server* server, int feed, size_t* block_bytes,
                          size_t* cb_bytes, size_t* ib_bytes);
/* Device-resident weights: copied from host memory (FBA char2) ... */
int fbf_server_set_weights(fbf_server* server, int feed, const void* weights_host, size_t bytes);
/* ... or generated on the device from delays ([B][A] float2, host) at
 * `epoch` for channel centres chan_freqs (host, nchans doubles).  Blocks
 * admitted after the call use the new weights; blocks already running keep
 * the old ones. */
int fbf_server_set_delays(fbf_server* server, int feed, const void* delays_host, size_t bytes,
                          const double* chan_freqs, double epoch);
int fbf_server_start(fbf_server* server);
/* Queues one block.  input: host (pinned for full speed) or device memory
 * (input_on_device != 0) of fbf_server_feed_bytes' block size; cb_out/ib_out:
 * host buffers for the beams (ib_out may be NULL without an incoherent beam).
 * Buffers must stay valid until the block completes.  Returns
 * FBF_ERR_CAPACITY if max_pending blocks are already queued. */
int fbf_server_submit(fbf_server* server, int feed, const void* input, size_t bytes, int input_on_device,
                      void* cb_out, void* ib_out, uint64_t tag);
/* Next completed block: waits up to timeout_ms (0 = no wait).  Returns
 * FBF_OK with *tag, *latency_ms and *status (FBF_OK or the block's error), or
 * FBF_ERR_STATE if nothing completed in time. */
int fbf_server_poll(fbf_server* server, uint64_t* tag, int* status, double* latency_ms, int timeout_ms);
int fbf_server_stats_get(fbf_server* server, fbf_server_stats* stats);
int fbf_server_reset_stats(fbf_server* server);
/* Waits until every submitted block has completed. */
int fbf_server_drain(fbf_server* server);

#ifdef __cplusplus
}
#endif
#endif /* FBFUSE_HIP_H */
