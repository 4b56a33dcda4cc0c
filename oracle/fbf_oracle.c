/*
 * fbf_oracle.c -- CPU restatement of the FBFUSE beamformer path.
 *
 * TEST INFRASTRUCTURE ONLY (see fbf_oracle.h): used by tests/, smoke() and
 * the cpu_baseline leg of bench.py, never by the product.
 *
 * Built with -ffp-contract=off so that the FP64 phase and quantiser
 * arithmetic is the plain IEEE sequence written here.
 */
#include "fbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int clamp_threads(int n)
{
    return n < 1 ? 1 : n;
}

/* ------------------------------------------------------------------ */
/* Split transpose: SplitTransposeTester.cu:33-75                      */
/* ------------------------------------------------------------------ */
void oracle_split_transpose(const int8_t* taftp, int8_t* ftpa,
                            int nheap_groups, int total_nantennas,
                            int used_nantennas, int start_antenna,
                            int nchans, int nsamples_per_heap, int npol)
{
    /* element = char2 (2 bytes) */
    const long tp = (long)nsamples_per_heap * npol;
    const long ftp = (long)nchans * tp;
    const long aftp = (long)total_nantennas * ftp;
    const long pa = (long)npol * used_nantennas;
    const long tpa = (long)nsamples_per_heap * nheap_groups * pa;
    for (long ts = 0; ts < nheap_groups; ++ts)
        for (long a = 0; a < used_nantennas; ++a) {
            long ia = a + start_antenna;
            for (long f = 0; f < nchans; ++f)
                for (long s = 0; s < nsamples_per_heap; ++s)
                    for (long p = 0; p < npol; ++p) {
                        long in_idx = ts * aftp + ia * ftp + f * tp + s * npol + p;
                        long out_s = ts * nsamples_per_heap + s;
                        long out_idx = f * tpa + out_s * pa + p * used_nantennas + a;
                        ftpa[2 * out_idx] = taftp[2 * in_idx];
                        ftpa[2 * out_idx + 1] = taftp[2 * in_idx + 1];
                    }
        }
}

/* ------------------------------------------------------------------ */
/* Coherent beam: real embedding of the complex weights.               */
/* Row Re_b = [wr_a, -wi_a], row Im_b = [wi_a, wr_a] over k = (a, re/im)*/
/* so that Re = sum wr*vr - wi*vi and Im = sum wi*vr + wr*vi, which is  */
/* r = xx - yy and i = xy + yx of CoherentBeamformer.cu:142-156 and     */
/* CoherentBeamformerTester.cu:88-93.  -wi is formed in int16, so it is */
/* exact for every int8 weight (128 for wi = -128), as the reference's  */
/* int32 dp4a arithmetic is.                                            */
/* ------------------------------------------------------------------ */

static void embed_channel(const int8_t* fba_f, int nbeams, int nantennas,
                          int16_t* wre, int16_t* wim)
{
    for (int b = 0; b < nbeams; ++b)
        for (int a = 0; a < nantennas; ++a) {
            int8_t wr = fba_f[2 * ((long)b * nantennas + a)];
            int8_t wi = fba_f[2 * ((long)b * nantennas + a) + 1];
            wre[(long)b * 2 * nantennas + 2 * a] = wr;
            wre[(long)b * 2 * nantennas + 2 * a + 1] = (int16_t)(-(int16_t)wi);
            wim[(long)b * 2 * nantennas + 2 * a] = wi;
            wim[(long)b * 2 * nantennas + 2 * a + 1] = wr;
        }
}

static inline int32_t dot16(const int16_t* x, const int16_t* y, int n)
{
    int32_t acc = 0;
    for (int k = 0; k < n; ++k)
        acc += (int32_t)x[k] * (int32_t)y[k];
    return acc;
}

void oracle_cb_sums(const int8_t* ftpa, const int8_t* fba_weights,
                    int32_t* sums, int nchans, long nsamples, int npol,
                    int nantennas, int nbeams, int nthreads)
{
    const int K = 2 * nantennas;
    for (int f = 0; f < nchans; ++f) {
        int16_t* wre = (int16_t*)malloc(sizeof(int16_t) * (size_t)nbeams * K);
        int16_t* wim = (int16_t*)malloc(sizeof(int16_t) * (size_t)nbeams * K);
        embed_channel(fba_weights + (long)f * nbeams * K, nbeams, nantennas, wre, wim);
        const long ncols = nsamples * npol;
#pragma omp parallel for num_threads(clamp_threads(nthreads)) schedule(static)
        for (long n = 0; n < ncols; ++n) {
            int16_t v[256 * 2];
            const int8_t* col = ftpa + ((long)f * ncols + n) * K;
            for (int k = 0; k < K; ++k)
                v[k] = col[k];
            int32_t* out = sums + ((long)f * ncols + n) * nbeams * 2;
            for (int b = 0; b < nbeams; ++b) {
                out[2 * b] = dot16(wre + (long)b * K, v, K);
                out[2 * b + 1] = dot16(wim + (long)b * K, v, K);
            }
        }
        free(wre);
        free(wim);
    }
}

void oracle_cb_power(const int8_t* ftpa, const int8_t* fba_weights,
                     double* power, int nchans, long nsamples, int npol,
                     int nantennas, int nbeams, int tscrunch, int fscrunch,
                     int nthreads)
{
    const int K = 2 * nantennas;
    const int nchans_out = nchans / fscrunch;
    const long nsamples_out = nsamples / tscrunch;
    const long ncols = nsamples * npol;
    int16_t* wre = (int16_t*)malloc(sizeof(int16_t) * (size_t)nchans * nbeams * K);
    int16_t* wim = (int16_t*)malloc(sizeof(int16_t) * (size_t)nchans * nbeams * K);
    for (int f = 0; f < nchans; ++f)
        embed_channel(fba_weights + (long)f * nbeams * K, nbeams, nantennas,
                      wre + (long)f * nbeams * K, wim + (long)f * nbeams * K);
#pragma omp parallel for collapse(2) num_threads(clamp_threads(nthreads)) schedule(dynamic, 1)
    for (int fo = 0; fo < nchans_out; ++fo)
        for (long to = 0; to < nsamples_out; ++to) {
            int64_t* acc = (int64_t*)calloc((size_t)nbeams, sizeof(int64_t));
            int16_t v[256 * 2];
            for (int fc = 0; fc < fscrunch; ++fc) {
                const int f = fo * fscrunch + fc;
                const int16_t* wr_f = wre + (long)f * nbeams * K;
                const int16_t* wi_f = wim + (long)f * nbeams * K;
                for (int ts = 0; ts < tscrunch; ++ts)
                    for (int p = 0; p < npol; ++p) {
                        long n = (to * tscrunch + ts) * npol + p;
                        const int8_t* col = ftpa + ((long)f * ncols + n) * K;
                        for (int k = 0; k < K; ++k)
                            v[k] = col[k];
                        for (int b = 0; b < nbeams; ++b) {
                            int64_t r = dot16(wr_f + (long)b * K, v, K);
                            int64_t i = dot16(wi_f + (long)b * K, v, K);
                            acc[b] += r * r + i * i;
                        }
                    }
            }
            double* out = power + ((long)fo * nsamples_out + to) * nbeams;
            for (int b = 0; b < nbeams; ++b)
                out[b] = (double)acc[b];
            free(acc);
        }
    free(wre);
    free(wim);
}

/* ------------------------------------------------------------------ */
/* Quantiser.  (P - offset) / scale in IEEE double, then round half     */
/* away from zero (ROUND) or truncate (TRUNC, the reference's cast at   */
/* CoherentBeamformer.cu:169), saturated to int8.                       */
/* ------------------------------------------------------------------ */
int8_t oracle_quantise(double power, float offset, float scale, int mode)
{
    volatile double num = power - (double)offset;
    volatile double v = num / (double)scale;
    double q = (mode == ORACLE_QUANT_TRUNC) ? trunc(v) : round(v);
    if (!(q >= -128.0)) /* also maps NaN to -128 */
        q = -128.0;
    if (q > 127.0)
        q = 127.0;
    return (int8_t)q;
}

void oracle_cb_tbtf(const double* power, int8_t* tbtf, const float* offsets,
                    const float* scales, int nbeams, int nchans_out,
                    long nsamples_out, int nsamples_per_heap_out, int mode)
{
    /* CoherentBeamformer.cu:162-168 / CoherentBeamformerTester.cu:101-107 */
    const long tf_size = (long)nsamples_per_heap_out * nchans_out;
    const long btf_size = (long)nbeams * tf_size;
    for (int fo = 0; fo < nchans_out; ++fo)
        for (long to = 0; to < nsamples_out; ++to)
            for (int b = 0; b < nbeams; ++b) {
                double p = power[((long)fo * nsamples_out + to) * nbeams + b];
                long idx = to / nsamples_per_heap_out * btf_size + (long)b * tf_size
                         + (to % nsamples_per_heap_out) * nchans_out + fo;
                tbtf[idx] = oracle_quantise(p, offsets[fo], scales[fo], mode);
            }
}

static int8_t sat_cast_trunc_f(float v)
{
    /* (int8_t) of a float: the reference relies on it being in range; the
     * restatement saturates so that it is defined everywhere. */
    float t = truncf(v);
    if (!(t >= -128.0f))
        return -128;
    if (t > 127.0f)
        return 127;
    return (int8_t)t;
}

void oracle_cb_reference_f32(const int8_t* ftpa, const int8_t* fba_weights,
                             int8_t* tbtf, int nchans, int tscrunch,
                             int fscrunch, long nsamples, int nbeams,
                             int nantennas, int npol, float scale,
                             float offset, int nsamples_per_heap_out)
{
    /* CoherentBeamformerTester.cu:37-118, loop for loop. */
#pragma omp parallel for schedule(dynamic, 1)
    for (int channel_idx = 0; channel_idx < nchans; channel_idx += fscrunch)
        for (long sample_idx = 0; sample_idx < nsamples; sample_idx += tscrunch)
            for (int beam_idx = 0; beam_idx < nbeams; ++beam_idx) {
                float power = 0.0f;
                for (int sc = channel_idx; sc < channel_idx + fscrunch; ++sc)
                    for (int so = 0; so < tscrunch; ++so)
                        for (int pol = 0; pol < npol; ++pol) {
                            float ax = 0.0f, ay = 0.0f;
                            for (int a = 0; a < nantennas; ++a) {
                                long vidx = (long)nantennas * npol * nsamples * sc
                                          + (long)nantennas * npol * (sample_idx + so)
                                          + (long)nantennas * pol + a;
                                long widx = (long)nantennas * nbeams * sc
                                          + (long)nantennas * beam_idx + a;
                                float dx = ftpa[2 * vidx], dy = ftpa[2 * vidx + 1];
                                float wx = fba_weights[2 * widx], wy = fba_weights[2 * widx + 1];
                                float xx = dx * wx, yy = dy * wy, xy = dx * wy, yx = dy * wx;
                                ax += xx - yy;
                                ay += xy + yx;
                            }
                            power += ax * ax + ay * ay;
                        }
                long tf_size = (long)nsamples_per_heap_out * nchans / fscrunch;
                long btf_size = (long)nbeams * tf_size;
                long osamp = sample_idx / tscrunch;
                long idx = osamp / nsamples_per_heap_out * btf_size + (long)beam_idx * tf_size
                         + (osamp % nsamples_per_heap_out) * (nchans / fscrunch)
                         + channel_idx / fscrunch;
                tbtf[idx] = sat_cast_trunc_f((power - offset) / scale);
            }
}

/* ------------------------------------------------------------------ */
/* Incoherent beam: IncoherentBeamformerTester.cu:36-97                 */
/* ------------------------------------------------------------------ */
void oracle_ib_power(const int8_t* taftp, double* power, int nchans,
                     int tscrunch, int fscrunch, int ntimestamps,
                     int nantennas, int npol, int nsamples_per_heap,
                     int nthreads)
{
    const long tp = (long)nsamples_per_heap * npol;
    const long ftp = (long)nchans * tp;
    const long aftp = (long)nantennas * ftp;
    const int nsub = nsamples_per_heap / tscrunch;
    const int nband = nchans / fscrunch;
#pragma omp parallel for collapse(2) num_threads(clamp_threads(nthreads))
    for (int t = 0; t < ntimestamps; ++t)
        for (int si = 0; si < nsub; ++si)
            for (int sb = 0; sb < nband; ++sb) {
                int64_t acc = 0;
                for (int a = 0; a < nantennas; ++a)
                    for (int c = sb * fscrunch; c < (sb + 1) * fscrunch; ++c)
                        for (int s = si * tscrunch; s < (si + 1) * tscrunch; ++s)
                            for (int p = 0; p < npol; ++p) {
                                long idx = t * aftp + a * ftp + c * tp + (long)s * npol + p;
                                int x = taftp[2 * idx], y = taftp[2 * idx + 1];
                                acc += x * x + y * y;
                            }
                long time_idx = (long)t * nsub + si;
                power[time_idx * nband + sb] = (double)acc;
            }
}

void oracle_ib_tf(const double* power, int8_t* tf, const float* offsets,
                  const float* scales, long n_time, int nchans_out, int mode)
{
    for (long t = 0; t < n_time; ++t)
        for (int c = 0; c < nchans_out; ++c)
            tf[t * nchans_out + c] = oracle_quantise(power[t * nchans_out + c],
                                                     offsets[c], scales[c], mode);
}

void oracle_ib_reference_f32(const int8_t* taftp, int8_t* tf, int nchans,
                             int tscrunch, int fscrunch, int ntimestamps,
                             int nantennas, int npol, int nsamples_per_heap,
                             float scale, float offset)
{
    const long tp = (long)nsamples_per_heap * npol;
    const long ftp = (long)nchans * tp;
    const long aftp = (long)nantennas * ftp;
    for (int t = 0; t < ntimestamps; ++t)
        for (int si = 0; si < nsamples_per_heap / tscrunch; ++si)
            for (int sb = 0; sb < nchans / fscrunch; ++sb) {
                float xx = 0.0f, yy = 0.0f;
                for (int a = 0; a < nantennas; ++a)
                    for (int c = sb * fscrunch; c < (sb + 1) * fscrunch; ++c)
                        for (int s = si * tscrunch; s < (si + 1) * tscrunch; ++s)
                            for (int p = 0; p < npol; ++p) {
                                long idx = t * aftp + a * ftp + c * tp + (long)s * npol + p;
                                float x = taftp[2 * idx], y = taftp[2 * idx + 1];
                                xx += x * x;
                                yy += y * y;
                            }
                long time_idx = (long)t * nsamples_per_heap / tscrunch + si;
                float pw = xx + yy;
                tf[time_idx * (nchans / fscrunch) + sb] = sat_cast_trunc_f((pw - offset) / scale);
            }
}

/* ------------------------------------------------------------------ */
/* Weights: WeightsManager.cu:58-73, WeightsManagerTester.cu:38-73      */
/* ------------------------------------------------------------------ */
/* Nearest integer, ties to even: __float2int_rn (WeightsManager.cu:70-71). */
static int8_t round_to_int8(double v)
{
    double r = rint(v);
    if (r > 127.0)
        r = 127.0;
    if (r < -128.0)
        r = -128.0;
    return (int8_t)r;
}

void oracle_weights_f64(const float* delays, const double* freqs,
                        double epoch, int nantennas, int nbeams, int nchans,
                        int8_t* weights, double* phase_frac)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int f = 0; f < nchans; ++f)
        for (int b = 0; b < nbeams; ++b)
            for (int a = 0; a < nantennas; ++a) {
                long di = (long)b * nantennas + a;
                double rate = (double)delays[2 * di];
                double off = (double)delays[2 * di + 1];
                volatile double tau = rate * epoch;
                tau = tau + off;
                volatile double phase = tau * freqs[f];
                double frac = phase - rint(phase);
                long wi = ((long)f * nbeams + b) * nantennas + a;
                weights[2 * wi] = round_to_int8(127.0 * cos(two_pi * frac));
                weights[2 * wi + 1] = round_to_int8(127.0 * sin(two_pi * frac));
                if (phase_frac)
                    phase_frac[wi] = frac;
            }
}

void oracle_weights_reference_f32(const float* delays, const float* freqs,
                                  float epoch, int nantennas, int nbeams,
                                  int nchans, int8_t* weights)
{
    const float twopi = 6.283185307179586f;
    for (int a = 0; a < nantennas; ++a)
        for (int b = 0; b < nbeams; ++b) {
            float rate = delays[2 * ((long)b * nantennas + a)];
            float off = delays[2 * ((long)b * nantennas + a) + 1];
            for (int c = 0; c < nchans; ++c) {
                volatile float prod = epoch * rate;
                volatile float sum = prod + off;
                volatile float phase = sum * freqs[c];
                volatile float arg = twopi * phase;
                float s = sinf(arg), co = cosf(arg);
                volatile float xr = co * 127.0f, xi = s * 127.0f;
                long oi = (long)nantennas * ((long)nbeams * c + b) + a;
                weights[2 * oi] = (int8_t)rintf(xr);
                weights[2 * oi + 1] = (int8_t)rintf(xi);
            }
        }
}

/* ------------------------------------------------------------------ */
/* PipelineConfig.cpp:109-127 and :180-195                              */
/* ------------------------------------------------------------------ */
void oracle_power_levels(float input_level, float output_level,
                         int cb_nantennas, int cb_tscrunch, int cb_fscrunch,
                         int ib_nantennas, int ib_tscrunch, int ib_fscrunch,
                         int npol, float* cb_offset, float* cb_scale,
                         float* ib_offset, float* ib_scale)
{
    const float weights_amp = 127.0f;
    volatile float amp = weights_amp * input_level * sqrtf((float)cb_nantennas);
    float cbs = (float)pow((double)amp, 2);
    float cb_dof = (float)(2 * (size_t)cb_tscrunch * cb_fscrunch * npol);
    *cb_offset = cbs * cb_dof;
    volatile float t = cbs * sqrtf(2 * cb_dof);
    *cb_scale = t / output_level;
    float ibs = (float)pow((double)input_level, 2);
    float ib_dof = (float)(2 * (size_t)ib_tscrunch * ib_fscrunch * ib_nantennas * npol);
    *ib_offset = ibs * ib_dof;
    volatile float u = ibs * sqrtf(2 * ib_dof);
    *ib_scale = u / output_level;
}

void oracle_channel_frequencies(float cfreq, float bw, int nchans, float* freqs)
{
    float chbw = bw / (float)nchans;
    float fbottom = cfreq - bw / 2.0f;
    for (int c = 0; c < nchans; ++c) {
        volatile float x = chbw * (float)c;
        volatile float y = fbottom + chbw / 2.0f;
        freqs[c] = y + x;
    }
}

/* ------------------------------------------------------------------ */
/* TUSE corner turn: meerkat/tuse/src/transpose_to_dada.cpp:19-48       */
/* ------------------------------------------------------------------ */
void oracle_tuse_transpose(const int8_t* in, int8_t* out, int nchans,
                           int nsamples, int nfreq, int nbeams, int ngroups,
                           int beam0, int nbeams_out)
{
    /* The reference copies one beam of one group (nfreq bands of
     * nsamples x nchans) and writes, for every sample and band, the band's
     * nchans bytes at isamp * nchans * nfreq + iband * nchans. */
    const size_t per_beam = (size_t)nchans * nsamples * nfreq;  /* skipbeam / beamnum */
    const size_t per_group = per_beam * nbeams;                  /* skipgroup */
    const size_t skipband = (size_t)nchans * nsamples;
    const size_t skipallchans = (size_t)nchans * nfreq;
    for (int bo = 0; bo < nbeams_out; ++bo) {
        const int8_t* inb = in + (size_t)(beam0 + bo) * per_beam;
        int8_t* outb = out + (size_t)bo * per_beam * ngroups;
        for (int g = 0; g < ngroups; ++g)
            for (int s = 0; s < nsamples; ++s)
                for (int f = 0; f < nfreq; ++f)
                    memcpy(outb + (size_t)g * per_beam + (size_t)s * skipallchans + (size_t)f * nchans,
                           inb + (size_t)g * per_group + (size_t)f * skipband + (size_t)s * nchans,
                           (size_t)nchans);
    }
}

/* ------------------------------------------------------------------ */
/* F-engine heaps: meerkat/tools/src/feng_to_dada.cu:54-80              */
/* ------------------------------------------------------------------ */
void oracle_feng_to_dada(const int32_t* in, int32_t* out, int nchans,
                         long ntimestamps)
{
    /* MEERKAT_FENG_NSAMPS_PER_HEAP = 256 (meerkat/constants.hpp:4); the
     * kernel's input index (chan * 256 + time) and output index
     * (time * nchans + chan) within each timestamp's block. */
    const long nsamps = 256;
    for (long t = 0; t < ntimestamps; ++t) {
        const int32_t* src = in + t * nchans * nsamps;
        int32_t* dst = out + t * nchans * nsamps;
        for (long s = 0; s < nsamps; ++s)
            for (long c = 0; c < nchans; ++c) dst[s * nchans + c] = src[c * nsamps + s];
    }
}
