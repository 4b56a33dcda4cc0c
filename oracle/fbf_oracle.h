/*
 * fbf_oracle.h -- CPU restatement of the MeerKAT FBFUSE beamformer path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * called only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product (psrdada_cpp_amd) never calls it and has no
 * CPU fallback.
 *
 * Every function names the reference file:line whose behaviour it restates
 * (paths relative to the psrdada_cpp source tree).  Where the library fixes a
 * semantic the reference leaves open, the choice is written next to the
 * function and mirrored exactly by the HIP kernels:
 *
 *   * int32 beam sums are exact (the reference's dp4a path is exact too).
 *   * detected power is accumulated exactly (int64, handed over as double;
 *     every partial sum is an integer below 2^53) instead of the
 *     reference's FP32 accumulator (CoherentBeamformer.cu:155-157).
 *   * the quantiser is (double(P) - double(offset)) / double(scale),
 *     correctly rounded, then FBF_QUANT_ROUND: round half away from zero,
 *     or FBF_QUANT_TRUNC: truncate toward zero (the reference's
 *     (int8_t) cast, CoherentBeamformer.cu:169), both saturated to
 *     [-128, 127].
 *   * the int32 beam sums are exact for every int8 weight, -128 included,
 *     as the reference's dp4a arithmetic is (CoherentBeamformer.cu:150-155).
 */
#ifndef FBF_ORACLE_H
#define FBF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_QUANT_ROUND = 0, ORACLE_QUANT_TRUNC = 1 };

/* SplitTransposeTester.cu:33-75 (and the kernel SplitTranspose.cu:11-70):
 * TAFTP [hg][A_tot][F][S][P] char2 -> FTPA [F][hg*S+s][P][A_used] char2. */
void oracle_split_transpose(const int8_t* taftp, int8_t* ftpa,
                            int nheap_groups, int total_nantennas,
                            int used_nantennas, int start_antenna,
                            int nchans, int nsamples_per_heap, int npol);

/* Exact complex beam sums, CoherentBeamformerTester.cu:75-96 /
 * CoherentBeamformer.cu:142-156 (no conjugation: r = wr*vr - wi*vi,
 * i = wr*vi + wi*vr).  sums: int32 [F][T][P][B][2] (re, im). */
void oracle_cb_sums(const int8_t* ftpa, const int8_t* fba_weights,
                    int32_t* sums, int nchans, long nsamples, int npol,
                    int nantennas, int nbeams, int nthreads);

/* Exact detected power, summed over pol, tscrunch and fscrunch
 * (CoherentBeamformerTester.cu:55-100).  power: double [F/fs][T/ts][B]. */
void oracle_cb_power(const int8_t* ftpa, const int8_t* fba_weights,
                     double* power, int nchans, long nsamples, int npol,
                     int nantennas, int nbeams, int tscrunch, int fscrunch,
                     int nthreads);

/* The quantiser (CoherentBeamformer.cu:169, PipelineConfig.hpp:186-199). */
int8_t oracle_quantise(double power, float offset, float scale, int mode);

/* Power [Fo][To][B] -> int8 TBTF [To/ohs][B][ohs][Fo] with per-output-channel
 * offset/scale (index map CoherentBeamformer.cu:162-168). */
void oracle_cb_tbtf(const double* power, int8_t* tbtf, const float* offsets,
                    const float* scales, int nbeams, int nchans_out,
                    long nsamples_out, int nsamples_per_heap_out, int mode);

/* Faithful FP32 restatement of beamformer_c_reference
 * (CoherentBeamformerTester.cu:37-118): float accumulators, float power,
 * truncating cast (saturated here, where the reference is undefined). */
void oracle_cb_reference_f32(const int8_t* ftpa, const int8_t* fba_weights,
                             int8_t* tbtf, int nchans, int tscrunch,
                             int fscrunch, long nsamples, int nbeams,
                             int nantennas, int npol, float scale,
                             float offset, int nsamples_per_heap_out);

/* Exact incoherent power, IncoherentBeamformerTester.cu:36-97:
 * P = sum over antennas, fscrunch channels, tscrunch samples and pols of
 * re^2 + im^2.  TAFTP input, power double [hg][S/ts][F/fs]. */
void oracle_ib_power(const int8_t* taftp, double* power, int nchans,
                     int tscrunch, int fscrunch, int ntimestamps,
                     int nantennas, int npol, int nsamples_per_heap,
                     int nthreads);

/* Power -> int8 TF (IncoherentBeamformerTester.cu:269-276), per-output-
 * channel offset/scale; n_time = ntimestamps * S / ts. */
void oracle_ib_tf(const double* power, int8_t* tf, const float* offsets,
                  const float* scales, long n_time, int nchans_out, int mode);

/* Faithful FP32 restatement of the IB host reference
 * (IncoherentBeamformerTester.cu:36-97). */
void oracle_ib_reference_f32(const int8_t* taftp, int8_t* tf, int nchans,
                             int tscrunch, int fscrunch, int ntimestamps,
                             int nantennas, int npol, int nsamples_per_heap,
                             float scale, float offset);

/* FP64 weights: phase = f * (rate * t + offset) (WeightsManager.cu:65-66,
 * delays .x = rate, .y = offset), w = round(127 * exp(i 2 pi phase)),
 * layout [F][B][A] char2 (WeightsManager.cu:73 with ntsteps = 1).
 * phase_frac (optional, [F][B][A]) receives phase - rint(phase). */
void oracle_weights_f64(const float* delays, const double* freqs,
                        double epoch, int nantennas, int nbeams, int nchans,
                        int8_t* weights, double* phase_frac);

/* Faithful FP32 restatement of calc_weights_c_reference
 * (WeightsManagerTester.cu:38-73). */
void oracle_weights_reference_f32(const float* delays, const float* freqs,
                                  float epoch, int nantennas, int nbeams,
                                  int nchans, int8_t* weights);

/* PipelineConfig.cpp:109-127, same float arithmetic. */
void oracle_power_levels(float input_level, float output_level,
                         int cb_nantennas, int cb_tscrunch, int cb_fscrunch,
                         int ib_nantennas, int ib_tscrunch, int ib_fscrunch,
                         int npol, float* cb_offset, float* cb_scale,
                         float* ib_offset, float* ib_scale);

/* PipelineConfig.cpp:180-195, same float arithmetic. */
void oracle_channel_frequencies(float cfreq, float bw, int nchans,
                                float* freqs);

/* TUSE beam corner turn (meerkat/tuse/src/transpose_to_dada.cpp:19-48),
 * do_transpose for each beam in [beam0, beam0 + nbeams_out):
 * in  [ngroups][nbeams][nfreq][nsamples][nchans] bytes,
 * out [nbeams_out][ngroups][nsamples][nfreq][nchans] bytes. */
void oracle_tuse_transpose(const int8_t* in, int8_t* out, int nchans,
                           int nsamples, int nfreq, int nbeams, int ngroups,
                           int beam0, int nbeams_out);

/* F-engine heaps to DADA order (meerkat/tools/src/feng_to_dada.cu:54-80):
 * per timestamp [nchans][256] 32-bit words (two char2 polarisations) ->
 * [256][nchans]. */
void oracle_feng_to_dada(const int32_t* in, int32_t* out, int nchans,
                         long ntimestamps);

#ifdef __cplusplus
}
#endif
#endif
