// Coherent beamformer on CDNA4 int8 MFMA.
//
// Behaviour follows CoherentBeamformer.cu:48-225 (kernel bf_aptf_general_k and
// CoherentBeamformer::beamform):
//   beam(b, f, t, p) = sum_a w[f][b][a] * v[f][t][p][a]      (no conjugate)
//   P(b, fo, to)     = sum over fscrunch, tscrunch, pol of |beam|^2
//   out              = quantise(P, offset[fo], scale[fo])   written TBTF
//
// MI355X design.  Per channel the contraction is a real GEMM
//   C[2B x T*P] = W'[2B x 2A] * V'[2A x T*P]
// with the [Re, -Im; Im, Re] embedding: row Re_b = [wr_a, -wi_a] and
// row Im_b = [wi_a, wr_a] over k = (antenna, re/im).  k is interleaved exactly
// as FTPA stores a column (char2 per antenna), so every 16-byte K slice of a
// voltage column is one contiguous 16-byte load straight into the MFMA B
// operand -- no de-interleave and no LDS staging.  The weights are embedded in
// registers from the raw FBA char2 bytes (byte swap for Im rows, saturating
// negation of the odd bytes for Re rows), once per channel per wave.
//
// v_mfma_i32_16x16x64_i8: A (16 rows x 64 k) = weights, B (64 k x 16 cols) =
// voltages, D lane l holds rows 4*(l/16)+i, column l%16.  Rows are ordered
// (beam, re/im) so one lane holds Re and Im of two beams for one column, and
// the 16 columns of a tile are 16 different output samples: column c of the
// j-th tile is input column (t0 + c) * tscrunch * P + j.  Detection and the
// whole tscrunch x pol x fscrunch integration therefore happen in-lane with no
// cross-lane traffic.  Power is accumulated in FP64, which is exact: every
// partial sum is an integer below 2^53 (|r| < 2^22 for A <= 128).
#include "fbf_internal.hpp"

namespace fbf {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

enum { MODE_INT8 = 0, MODE_SUMS = 1, MODE_POWER = 2 };

struct CbArgs {
    const int8_t* ftpa;
    const int8_t* fba;
    void* out;
    const float* offsets;
    const float* scales;
    long long nsamples;   // T
    long long nsamples_out;
    int A, B, F, Fo, ts, fs, ohs, quant;
    int n_bchunks, n_tchunks, n_bgroups, n_tgroups, wb;
};

// 16 bytes of one K slice [k0, k0 + 16) of a column of K bytes; bytes at
// k >= K read as zero.  K is a multiple of 8.
__device__ __forceinline__ v4i load_kslice(const int8_t* col, int k0, int K, bool al16)
{
    v4i r = {0, 0, 0, 0};
    if (k0 + 16 <= K) {
        if (al16) {
            r = *reinterpret_cast<const v4i*>(col + k0);
        } else {
            const int2 lo = *reinterpret_cast<const int2*>(col + k0);
            const int2 hi = *reinterpret_cast<const int2*>(col + k0 + 8);
            r = v4i{lo.x, lo.y, hi.x, hi.y};
        }
    } else if (k0 + 8 <= K) {
        const int2 lo = *reinterpret_cast<const int2*>(col + k0);
        r = v4i{lo.x, lo.y, 0, 0};
    }
    return r;
}

// Im row: [wr, wi] -> [wi, wr] per antenna (swap bytes in each 16-bit half).
__device__ __forceinline__ uint32_t swap_pairs(uint32_t x)
{
    return ((x >> 8) & 0x00FF00FFu) | ((x << 8) & 0xFF00FF00u);
}

__device__ __forceinline__ uint32_t sat_neg_byte(uint32_t b)
{
    const int v = (int)(int8_t)(b & 0xFF);
    return (uint32_t)(v == -128 ? 127 : -v) & 0xFFu;
}

// Re row: [wr, wi] -> [wr, -wi] per antenna, negation saturating at 127.
__device__ __forceinline__ uint32_t neg_odd(uint32_t x)
{
    return (x & 0x00FF00FFu) | (sat_neg_byte(x >> 8) << 8) | (sat_neg_byte(x >> 24) << 24);
}

template <int ROWT, int KSTEPS, int MODE>
__global__ __launch_bounds__(256) void cb_mfma_k(CbArgs p)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;   // K group of the operands, row group of D
    const int c = lane & 15;   // A row / B,D column
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int bgroup = logical % p.n_bgroups;
    const int rest = logical / p.n_bgroups;
    const int tgroup = rest % p.n_tgroups;
    const int fo = rest / p.n_tgroups;
    const int wt = 4 / p.wb;
    const int bc = bgroup * p.wb + (wave % p.wb);
    const int tc = tgroup * wt + (wave / p.wb);
    if (bc >= p.n_bchunks || tc >= p.n_tchunks || fo >= p.Fo) return;

    const int K = 2 * p.A;
    const bool al16 = (K & 15) == 0;
    const int beam_base = bc * (8 * ROWT);
    const long long t_out = (long long)tc * 16 + c;
    const bool t_valid = t_out < p.nsamples_out;
    const int tsp = 2 * p.ts;
    const long long TP = 2 * p.nsamples;

    double pw[ROWT][2];
#pragma unroll
    for (int r = 0; r < ROWT; ++r) pw[r][0] = pw[r][1] = 0.0;

    for (int fc = 0; fc < p.fs; ++fc) {
        const int f = fo * p.fs + fc;
        // Embedded weights for this channel: lane holds row c of each 16-row tile.
        v4i w[ROWT][KSTEPS];
#pragma unroll
        for (int r = 0; r < ROWT; ++r) {
            const int b = beam_base + r * 8 + (c >> 1);
            const int8_t* wcol = p.fba + ((size_t)f * p.B + b) * K;
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) {
                v4i raw = {0, 0, 0, 0};
                if (b < p.B) raw = load_kslice(wcol, s * 64 + g * 16, K, al16);
                v4i e;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    e[q] = (int)((c & 1) ? swap_pairs((uint32_t)raw[q]) : neg_odd((uint32_t)raw[q]));
                w[r][s] = e;
            }
        }

        const int8_t* vcol = p.ftpa + ((size_t)f * TP + (size_t)t_out * tsp) * K;
        v4i v[KSTEPS];
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s)
            v[s] = t_valid ? load_kslice(vcol, s * 64 + g * 16, K, al16) : v4i{0, 0, 0, 0};

        for (int j = 0; j < tsp; ++j) {
            // Prefetch the next column while this one is multiplied.
            v4i vn[KSTEPS];
            const bool more = (j + 1 < tsp) && t_valid;
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s)
                vn[s] = more ? load_kslice(vcol + (size_t)(j + 1) * K, s * 64 + g * 16, K, al16)
                             : v4i{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < ROWT; ++r) {
                v4i acc = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KSTEPS; ++s)
                    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(w[r][s], v[s], acc, 0, 0, 0);
                if (MODE == MODE_SUMS) {
                    const int b0 = beam_base + r * 8 + 2 * g;
                    if (t_valid && b0 < p.B) {
                        const long long n = t_out * tsp + j;
                        int32_t* dst = static_cast<int32_t*>(p.out) +
                                       (((size_t)f * TP + (size_t)n) * p.B + b0) * 2;
                        *reinterpret_cast<v4i*>(dst) = acc;
                    }
                }
                const double x0 = (double)acc[0], x1 = (double)acc[1];
                const double x2 = (double)acc[2], x3 = (double)acc[3];
                pw[r][0] += x0 * x0 + x1 * x1;
                pw[r][1] += x2 * x2 + x3 * x3;
            }
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) v[s] = vn[s];
        }
    }

    if (MODE == MODE_SUMS || !t_valid) return;
    if (MODE == MODE_POWER) {
        double* dst = static_cast<double*>(p.out) + ((size_t)fo * p.nsamples_out + t_out) * p.B;
#pragma unroll
        for (int r = 0; r < ROWT; ++r) {
            const int b0 = beam_base + r * 8 + 2 * g;
            if (b0 < p.B) {
                dst[b0] = pw[r][0];
                dst[b0 + 1] = pw[r][1];
            }
        }
        return;
    }
    // int8 TBTF: [to / ohs][b][to % ohs][fo]
    const float off = p.offsets[fo];
    const float sc = p.scales[fo];
    int8_t* out = static_cast<int8_t*>(p.out);
    const size_t tf_size = (size_t)p.ohs * p.Fo;
    const size_t heap_base = (size_t)(t_out / p.ohs) * p.B * tf_size + (size_t)(t_out % p.ohs) * p.Fo + fo;
#pragma unroll
    for (int r = 0; r < ROWT; ++r) {
        const int b0 = beam_base + r * 8 + 2 * g;
        if (b0 < p.B) {
            out[heap_base + (size_t)b0 * tf_size] = quantise(pw[r][0], off, sc, p.quant);
            out[heap_base + (size_t)(b0 + 1) * tf_size] = quantise(pw[r][1], off, sc, p.quant);
        }
    }
}

template <int ROWT, int MODE>
hipError_t launch_rowt(const CbArgs& a, int ksteps, dim3 grid, hipStream_t s)
{
    switch (ksteps) {
    case 1: cb_mfma_k<ROWT, 1, MODE><<<grid, 256, 0, s>>>(a); break;
    case 2: cb_mfma_k<ROWT, 2, MODE><<<grid, 256, 0, s>>>(a); break;
    case 3: cb_mfma_k<ROWT, 3, MODE><<<grid, 256, 0, s>>>(a); break;
    case 4: cb_mfma_k<ROWT, 4, MODE><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const CbArgs& a, int rowt, int ksteps, dim3 grid, hipStream_t s)
{
    return rowt == 8 ? launch_rowt<8, MODE>(a, ksteps, grid, s) : launch_rowt<4, MODE>(a, ksteps, grid, s);
}

}  // namespace
}  // namespace fbf

// ---------------------------------------------------------------------------
// Plan and C ABI
// ---------------------------------------------------------------------------
using namespace fbf;

struct fbf_cb_plan {
    fbf_cb_desc d;
    int ohs;
    int variant;
    float* d_levels;  // device: offsets[Fo] then scales[Fo]
};

static int cb_validate(const fbf_cb_desc* d)
{
    FBF_REQUIRE(d, FBF_ERR_ARG, "fbf_cb: null descriptor");
    FBF_REQUIRE(d->nantennas > 0 && d->nbeams > 0 && d->nchans > 0 && d->tscrunch > 0 && d->fscrunch > 0,
                FBF_ERR_ARG, "fbf_cb: geometry values must be positive");
    FBF_REQUIRE(d->nantennas % 4 == 0 && d->nantennas <= 128, FBF_ERR_UNSUPPORTED,
                "fbf_cb: nantennas must be a multiple of 4 and <= 128 (CoherentBeamformer.cu:60-70)");
    FBF_REQUIRE(d->nbeams % 8 == 0, FBF_ERR_UNSUPPORTED, "fbf_cb: nbeams must be a multiple of 8");
    FBF_REQUIRE(d->nchans % d->fscrunch == 0, FBF_ERR_ARG, "fbf_cb: fscrunch must divide nchans");
    FBF_REQUIRE(d->quant_mode == FBF_QUANT_ROUND || d->quant_mode == FBF_QUANT_TRUNC, FBF_ERR_ARG,
                "fbf_cb: bad quant_mode");
    FBF_REQUIRE(d->nsamples_per_heap_out >= 0, FBF_ERR_ARG, "fbf_cb: bad nsamples_per_heap_out");
    return FBF_OK;
}

extern "C" int fbf_cb_plan_create(const fbf_cb_desc* desc, const float* offsets, const float* scales,
                                  fbf_cb_plan** plan)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb_plan_create: null plan pointer");
    *plan = nullptr;
    int st = cb_validate(desc);
    if (st) return st;
    FBF_REQUIRE(offsets && scales, FBF_ERR_ARG, "fbf_cb_plan_create: null offsets/scales");
    const int Fo = desc->nchans / desc->fscrunch;
    int ohs = desc->nsamples_per_heap_out ? desc->nsamples_per_heap_out : 8192 / Fo;
    FBF_REQUIRE(ohs > 0, FBF_ERR_ARG, "fbf_cb_plan_create: output heap holds no samples");
    auto* p = new fbf_cb_plan{*desc, ohs, 0, nullptr};
    hipError_t e = hipMalloc(&p->d_levels, sizeof(float) * 2 * Fo);
    if (e != hipSuccess) {
        delete p;
        return hip_status(e, "hipMalloc(levels)");
    }
    st = fbf_cb_plan_set_levels(p, offsets, scales);
    if (st) {
        hipFree(p->d_levels);
        delete p;
        return st;
    }
    *plan = p;
    return FBF_OK;
}

extern "C" int fbf_cb_plan_set_levels(fbf_cb_plan* plan, const float* offsets, const float* scales)
{
    FBF_REQUIRE(plan && offsets && scales, FBF_ERR_ARG, "fbf_cb_plan_set_levels: null argument");
    const int Fo = plan->d.nchans / plan->d.fscrunch;
    FBF_HIP_TRY(hipMemcpy(plan->d_levels, offsets, sizeof(float) * Fo, hipMemcpyHostToDevice));
    FBF_HIP_TRY(hipMemcpy(plan->d_levels + Fo, scales, sizeof(float) * Fo, hipMemcpyHostToDevice));
    return FBF_OK;
}

extern "C" int fbf_cb_plan_destroy(fbf_cb_plan* plan)
{
    if (!plan) return FBF_OK;
    hipError_t e = hipFree(plan->d_levels);
    delete plan;
    if (e != hipSuccess) return hip_status(e, "hipFree(levels)");
    return FBF_OK;
}

// Variant table: rows per wave tile.  0 = automatic.
static const int kCbVariantRowt[] = {0, 8, 4};

extern "C" int fbf_cb_variant_count(int* count)
{
    FBF_REQUIRE(count, FBF_ERR_ARG, "fbf_cb_variant_count: null pointer");
    *count = (int)(sizeof(kCbVariantRowt) / sizeof(kCbVariantRowt[0]));
    return FBF_OK;
}

extern "C" int fbf_cb_plan_set_variant(fbf_cb_plan* plan, int variant)
{
    int n = 0;
    fbf_cb_variant_count(&n);
    FBF_REQUIRE(plan && variant >= 0 && variant < n, FBF_ERR_ARG, "fbf_cb_plan_set_variant: bad variant");
    plan->variant = variant;
    return FBF_OK;
}

static int cb_nsamples(const fbf_cb_plan* plan, size_t ftpa_bytes, long long* nsamples)
{
    const size_t per_sample = (size_t)2 * 2 * plan->d.nantennas * plan->d.nchans;
    FBF_REQUIRE(ftpa_bytes % per_sample == 0, FBF_ERR_SIZE,
                "fbf_cb: input is not a whole number of samples (CoherentBeamformer.cu:197)");
    const long long T = (long long)(ftpa_bytes / per_sample);
    FBF_REQUIRE(T % plan->d.tscrunch == 0, FBF_ERR_SIZE, "fbf_cb: nsamples is not a multiple of tscrunch");
    *nsamples = T;
    return FBF_OK;
}

extern "C" int fbf_cb_output_bytes(const fbf_cb_plan* plan, size_t ftpa_bytes, size_t* tbtf_bytes)
{
    FBF_REQUIRE(plan && tbtf_bytes, FBF_ERR_ARG, "fbf_cb_output_bytes: null argument");
    long long T = 0;
    int st = cb_nsamples(plan, ftpa_bytes, &T);
    if (st) return st;
    *tbtf_bytes = (size_t)(T / plan->d.tscrunch) * plan->d.nbeams * (plan->d.nchans / plan->d.fscrunch);
    return FBF_OK;
}

static int cb_run(fbf_cb_plan* plan, int mode, const void* ftpa, size_t ftpa_bytes, const void* fba,
                  size_t weights_bytes, void* out, size_t out_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb: null plan");
    FBF_REQUIRE(ftpa && fba && out, FBF_ERR_ARG, "fbf_cb: null buffer");
    const fbf_cb_desc& d = plan->d;
    long long T = 0;
    int st = cb_nsamples(plan, ftpa_bytes, &T);
    if (st) return st;
    const size_t wbytes = (size_t)2 * d.nchans * d.nbeams * d.nantennas;
    FBF_REQUIRE(weights_bytes == wbytes, FBF_ERR_SIZE,
                "fbf_cb: weights size " + std::to_string(weights_bytes) + " != expected " +
                    std::to_string(wbytes) + " (CoherentBeamformer.cu:209)");
    const int Fo = d.nchans / d.fscrunch;
    const long long To = T / d.tscrunch;
    size_t expect = 0;
    if (mode == MODE_INT8) {
        FBF_REQUIRE(To % plan->ohs == 0, FBF_ERR_SIZE,
                    "fbf_cb: output samples " + std::to_string(To) + " are not a multiple of the output heap (" +
                        std::to_string(plan->ohs) + ") (CoherentBeamformer.cu:204)");
        expect = (size_t)To * d.nbeams * Fo;
    } else if (mode == MODE_SUMS) {
        expect = (size_t)d.nchans * T * 2 * d.nbeams * 2 * sizeof(int32_t);
    } else {
        expect = (size_t)Fo * To * d.nbeams * sizeof(double);
    }
    FBF_REQUIRE(out_bytes == expect, FBF_ERR_SIZE,
                "fbf_cb: output size " + std::to_string(out_bytes) + " != expected " + std::to_string(expect));
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(ftpa) & 15) == 0 && (reinterpret_cast<uintptr_t>(fba) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(out) & 15) == 0,
                FBF_ERR_ARG, "fbf_cb: buffers must be 16-byte aligned");
    if (T == 0) return FBF_OK;

    int rowt = kCbVariantRowt[plan->variant];
    if (rowt == 0) rowt = d.nbeams >= 64 ? 8 : 4;
    const int ksteps = (2 * d.nantennas + 63) / 64;
    CbArgs a;
    a.ftpa = static_cast<const int8_t*>(ftpa);
    a.fba = static_cast<const int8_t*>(fba);
    a.out = out;
    a.offsets = plan->d_levels;
    a.scales = plan->d_levels + Fo;
    a.nsamples = T;
    a.nsamples_out = To;
    a.A = d.nantennas;
    a.B = d.nbeams;
    a.F = d.nchans;
    a.Fo = Fo;
    a.ts = d.tscrunch;
    a.fs = d.fscrunch;
    a.ohs = plan->ohs;
    a.quant = d.quant_mode;
    a.n_bchunks = (d.nbeams + 8 * rowt - 1) / (8 * rowt);
    a.n_tchunks = (int)((To + 15) / 16);
    a.wb = a.n_bchunks >= 4 ? 4 : (a.n_bchunks >= 2 ? 2 : 1);
    a.n_bgroups = (a.n_bchunks + a.wb - 1) / a.wb;
    const int wt = 4 / a.wb;
    a.n_tgroups = (a.n_tchunks + wt - 1) / wt;
    const long long nblocks = (long long)Fo * a.n_tgroups * a.n_bgroups;
    FBF_REQUIRE(nblocks > 0 && nblocks < (1ll << 31), FBF_ERR_UNSUPPORTED, "fbf_cb: grid too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    dim3 grid((unsigned)nblocks);
    hipError_t e;
    if (mode == MODE_INT8)
        e = launch_mode<MODE_INT8>(a, rowt, ksteps, grid, s);
    else if (mode == MODE_SUMS)
        e = launch_mode<MODE_SUMS>(a, rowt, ksteps, grid, s);
    else
        e = launch_mode<MODE_POWER>(a, rowt, ksteps, grid, s);
    if (e != hipSuccess) return hip_status(e, "cb_mfma_k launch");
    return FBF_OK;
}

extern "C" int fbf_cb_beamform(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes, const void* fba_weights,
                               size_t weights_bytes, void* tbtf, size_t tbtf_bytes, fbf_stream_t stream)
{
    return cb_run(plan, MODE_INT8, ftpa, ftpa_bytes, fba_weights, weights_bytes, tbtf, tbtf_bytes, stream);
}

extern "C" int fbf_cb_beam_sums(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes, const void* fba_weights,
                                size_t weights_bytes, int32_t* sums, size_t sums_bytes, fbf_stream_t stream)
{
    return cb_run(plan, MODE_SUMS, ftpa, ftpa_bytes, fba_weights, weights_bytes, sums, sums_bytes, stream);
}

extern "C" int fbf_cb_beam_power(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes, const void* fba_weights,
                                 size_t weights_bytes, double* power, size_t power_bytes, fbf_stream_t stream)
{
    return cb_run(plan, MODE_POWER, ftpa, ftpa_bytes, fba_weights, weights_bytes, power, power_bytes, stream);
}
