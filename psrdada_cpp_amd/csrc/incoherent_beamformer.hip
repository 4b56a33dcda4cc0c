// Incoherent beam: detected power summed over all antennas.
//
// Behaviour follows IncoherentBeamformer.cu:10-117 (icbf_taftp_general_k and
// IncoherentBeamformer::beamform) and its host reference
// IncoherentBeamformerTester.cu:36-97:
//   P(to, fo) = sum over antennas, fscrunch channels, tscrunch samples and
//               pols of re^2 + im^2,  out = quantise(P, offset[fo], scale[fo])
// written TF: [heap_group * S / ts + window][fo].
//
// MI355X design: HBM-bound read of the whole TAFTP block.  A 256-thread block
// owns one heap group and `cpb` channels; each 64-thread row reads one
// 1 KiB (channel, antenna) row as 16-byte vectors (4 samples x 2 pols), so
// every load instruction is fully coalesced, and squares with v_dot4_i32_i8
// (exact int32).  tscrunch is reduced in-lane and across lanes with
// shuffles, fscrunch across rows through LDS, in int64 (exact).
//
// fbf_ib_beamform_weights also generates the block's beamforming weights in
// the same launch (its first work-groups), as the split-transpose launch can.
#include "fbf_internal.hpp"

namespace fbf {
namespace {

enum { IB_INT8 = 0, IB_POWER = 1 };

struct IbArgs {
    const int8_t* taftp;
    void* out;
    const float* offsets;
    const float* scales;
    int nhg, A, F, Fo, ts, fs, quant, cpb, nfb;
};

__device__ __forceinline__ int sq4(int x, int acc)
{
    return __builtin_amdgcn_sdot4(x, x, acc, false);
}

template <int MODE>
__global__ __launch_bounds__(256) void ib_k(IbArgs p, WeightsJob wj)
{
    constexpr int S = 256;
    __shared__ long long part[4][S];
    const int tid = threadIdx.x;
    if (weights_job_run(wj, tid)) return;   // uniform over the work-group
    const int sc = tid & 63;   // 16-byte chunk: samples 4*sc .. 4*sc+3
    const int row = tid >> 6;  // 0..3
    const int logical = xcd_remap((int)blockIdx.x - wj.nblocks, (int)gridDim.x - wj.nblocks);
    const int fb = logical % p.nfb;
    const int hg = logical / p.nfb;
    if (hg >= p.nhg) return;

    // One word squared is at most 4 * 128^2 = 2^16, so the int32 partials
    // over one channel's antennas are exact for A < 2^15 (fbf_ib_plan_create
    // checks it); each channel's partials are then added in int64.
    long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    const size_t row_bytes = (size_t)S * 4;
    for (int k = 0; k < p.cpb / 4; ++k) {
        int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        const int f = fb * p.cpb + row + 4 * k;
        const int8_t* base = p.taftp + ((size_t)hg * p.A * p.F + f) * row_bytes + (size_t)sc * 16;
        const size_t astride = (size_t)p.F * row_bytes;
        int a = 0;
        for (; a + 4 <= p.A; a += 4) {
            int4 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = *reinterpret_cast<const int4*>(base + (size_t)(a + u) * astride);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                c0 = sq4(x[u].x, c0);
                c1 = sq4(x[u].y, c1);
                c2 = sq4(x[u].z, c2);
                c3 = sq4(x[u].w, c3);
            }
        }
        for (; a < p.A; ++a) {
            const int4 x = *reinterpret_cast<const int4*>(base + (size_t)a * astride);
            c0 = sq4(x.x, c0);
            c1 = sq4(x.y, c1);
            c2 = sq4(x.z, c2);
            c3 = sq4(x.w, c3);
        }
        s0 += c0;
        s1 += c1;
        s2 += c2;
        s3 += c3;
    }

    // tscrunch: per-sample sums s0..s3 cover samples 4*sc + {0..3}.
    const int nwin = S / p.ts;
    if (p.ts >= 4) {
        long long v = s0 + s1 + s2 + s3;
        for (int m = 1; m < p.ts / 4; m <<= 1) v += __shfl_xor(v, m, 64);
        if ((sc % (p.ts / 4)) == 0) part[row][sc / (p.ts / 4)] = v;
    } else if (p.ts == 2) {
        part[row][2 * sc] = s0 + s1;
        part[row][2 * sc + 1] = s2 + s3;
    } else {
        part[row][4 * sc] = s0;
        part[row][4 * sc + 1] = s1;
        part[row][4 * sc + 2] = s2;
        part[row][4 * sc + 3] = s3;
    }
    __syncthreads();

    // fscrunch: rows belonging to one output channel are summed.
    const int rows_per_fo = p.fs >= 4 ? 4 : p.fs;
    const int nfo_blk = 4 / rows_per_fo;
    for (int idx = tid; idx < nwin * nfo_blk; idx += 256) {
        const int fl = idx / nwin;
        const int w = idx - fl * nwin;
        long long v = 0;
        for (int r = 0; r < rows_per_fo; ++r) v += part[fl * rows_per_fo + r][w];
        const int fo = (fb * p.cpb) / p.fs + fl;
        const size_t t = (size_t)hg * nwin + w;
        if (MODE == IB_POWER) {
            static_cast<double*>(p.out)[t * p.Fo + fo] = (double)v;
        } else {
            static_cast<int8_t*>(p.out)[t * p.Fo + fo] = quantise((double)v, p.offsets[fo], p.scales[fo], p.quant);
        }
    }
}

}  // namespace
}  // namespace fbf

using namespace fbf;

struct fbf_ib_plan {
    fbf_ib_desc d;
    float* d_levels;
};

extern "C" int fbf_ib_plan_create(const fbf_ib_desc* desc, const float* offsets, const float* scales,
                                  fbf_ib_plan** plan)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_ib_plan_create: null plan pointer");
    *plan = nullptr;
    FBF_REQUIRE(desc && offsets && scales, FBF_ERR_ARG, "fbf_ib_plan_create: null argument");
    const fbf_ib_desc& d = *desc;
    FBF_REQUIRE(d.nantennas > 0 && d.nchans > 0 && d.tscrunch > 0 && d.fscrunch > 0, FBF_ERR_ARG,
                "fbf_ib: geometry values must be positive");
    // Per-channel partial power sums are int32: a word (one sample, both
    // pols) squares to at most 2^16, so A antennas stay below 2^31 for A < 2^15.
    FBF_REQUIRE(d.nantennas < 32768, FBF_ERR_UNSUPPORTED,
                "fbf_ib: nantennas must be below 32768 (int32 per-channel partial power sums)");
    FBF_REQUIRE(d.nsamples_per_heap == 256, FBF_ERR_UNSUPPORTED,
                "fbf_ib: nsamples_per_heap must be 256 (FBFUSE_NSAMPLES_PER_HEAP)");
    FBF_REQUIRE(256 % d.tscrunch == 0, FBF_ERR_UNSUPPORTED, "fbf_ib: tscrunch must divide 256");
    FBF_REQUIRE(d.nchans % d.fscrunch == 0, FBF_ERR_ARG, "fbf_ib: fscrunch must divide nchans");
    FBF_REQUIRE((d.fscrunch <= 4 && 4 % d.fscrunch == 0) || d.fscrunch % 4 == 0, FBF_ERR_UNSUPPORTED,
                "fbf_ib: fscrunch must be 1, 2 or a multiple of 4");
    const int cpb = d.fscrunch > 4 ? d.fscrunch : 4;
    FBF_REQUIRE(d.nchans % cpb == 0, FBF_ERR_UNSUPPORTED, "fbf_ib: nchans must be a multiple of 4 and of fscrunch");
    FBF_REQUIRE(d.quant_mode == FBF_QUANT_ROUND || d.quant_mode == FBF_QUANT_TRUNC, FBF_ERR_ARG,
                "fbf_ib: bad quant_mode");
    auto* p = new fbf_ib_plan{d, nullptr};
    const int Fo = d.nchans / d.fscrunch;
    hipError_t e = hipMalloc(&p->d_levels, sizeof(float) * 2 * Fo);
    if (e != hipSuccess) {
        delete p;
        return hip_status(e, "hipMalloc(levels)");
    }
    int st = fbf_ib_plan_set_levels(p, offsets, scales);
    if (st) {
        (void)hipFree(p->d_levels);
        delete p;
        return st;
    }
    *plan = p;
    return FBF_OK;
}

extern "C" int fbf_ib_plan_set_levels(fbf_ib_plan* plan, const float* offsets, const float* scales)
{
    FBF_REQUIRE(plan && offsets && scales, FBF_ERR_ARG, "fbf_ib_plan_set_levels: null argument");
    const int Fo = plan->d.nchans / plan->d.fscrunch;
    // Work still running on any stream may read the levels being replaced.
    FBF_HIP_TRY(hipDeviceSynchronize());
    FBF_HIP_TRY(hipMemcpy(plan->d_levels, offsets, sizeof(float) * Fo, hipMemcpyHostToDevice));
    FBF_HIP_TRY(hipMemcpy(plan->d_levels + Fo, scales, sizeof(float) * Fo, hipMemcpyHostToDevice));
    return FBF_OK;
}

extern "C" int fbf_ib_plan_destroy(fbf_ib_plan* plan)
{
    if (!plan) return FBF_OK;
    hipError_t e = hipFree(plan->d_levels);
    delete plan;
    if (e != hipSuccess) return hip_status(e, "hipFree(levels)");
    return FBF_OK;
}

static int ib_nhg(const fbf_ib_plan* plan, size_t taftp_bytes, long long* nhg)
{
    const size_t hg_bytes = (size_t)4 * plan->d.nantennas * plan->d.nchans * plan->d.nsamples_per_heap;
    FBF_REQUIRE(taftp_bytes % hg_bytes == 0, FBF_ERR_SIZE,
                "fbf_ib: input is not a multiple of AFTP blocks (IncoherentBeamformer.cu:95)");
    *nhg = (long long)(taftp_bytes / hg_bytes);
    return FBF_OK;
}

extern "C" int fbf_ib_output_bytes(const fbf_ib_plan* plan, size_t taftp_bytes, size_t* tf_bytes)
{
    FBF_REQUIRE(plan && tf_bytes, FBF_ERR_ARG, "fbf_ib_output_bytes: null argument");
    long long nhg = 0;
    int st = ib_nhg(plan, taftp_bytes, &nhg);
    if (st) return st;
    *tf_bytes = (size_t)nhg * (plan->d.nsamples_per_heap / plan->d.tscrunch) * (plan->d.nchans / plan->d.fscrunch);
    return FBF_OK;
}

static int ib_run(fbf_ib_plan* plan, int mode, const void* taftp, size_t taftp_bytes, void* out, size_t out_bytes,
                  fbf_stream_t stream, WeightsJob wj = WeightsJob{})
{
    FBF_REQUIRE(plan && taftp && out, FBF_ERR_ARG, "fbf_ib: null argument");
    long long nhg = 0;
    int st = ib_nhg(plan, taftp_bytes, &nhg);
    if (st) return st;
    const fbf_ib_desc& d = plan->d;
    const int Fo = d.nchans / d.fscrunch;
    const size_t nout = (size_t)nhg * (d.nsamples_per_heap / d.tscrunch) * Fo;
    const size_t expect = mode == IB_INT8 ? nout : nout * sizeof(double);
    FBF_REQUIRE(out_bytes == expect, FBF_ERR_SIZE,
                "fbf_ib: output size " + std::to_string(out_bytes) + " != expected " + std::to_string(expect));
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(taftp) & 15) == 0, FBF_ERR_ARG, "fbf_ib: input must be 16-byte aligned");
    if (nhg == 0 && wj.nblocks == 0) return FBF_OK;
    IbArgs a;
    a.taftp = static_cast<const int8_t*>(taftp);
    a.out = out;
    a.offsets = plan->d_levels;
    a.scales = plan->d_levels + Fo;
    a.nhg = (int)nhg;
    a.A = d.nantennas;
    a.F = d.nchans;
    a.Fo = Fo;
    a.ts = d.tscrunch;
    a.fs = d.fscrunch;
    a.quant = d.quant_mode;
    a.cpb = d.fscrunch > 4 ? d.fscrunch : 4;
    a.nfb = d.nchans / a.cpb;
    const long long nblocks = nhg * a.nfb + wj.nblocks;
    FBF_REQUIRE(nblocks < (1ll << 31), FBF_ERR_UNSUPPORTED, "fbf_ib: grid too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (mode == IB_INT8)
        ib_k<IB_INT8><<<dim3((unsigned)nblocks), 256, 0, s>>>(a, wj);
    else
        ib_k<IB_POWER><<<dim3((unsigned)nblocks), 256, 0, s>>>(a, wj);
    FBF_LAUNCH_CHECK("ib_k");
    return FBF_OK;
}

int fbf::ib_finish_job(const fbf_ib_plan* plan, const int* partials, long long T, void* tf, size_t tf_bytes,
                       IbFinishJob* out)
{
    FBF_REQUIRE(plan && partials && tf && out, FBF_ERR_ARG, "fbf_ib: null argument");
    const fbf_ib_desc& d = plan->d;
    FBF_REQUIRE(T % 16 == 0 && T % d.tscrunch == 0, FBF_ERR_SIZE, "fbf_ib: partials of a partial tile");
    FBF_REQUIRE(d.tscrunch < 16 ? 16 % d.tscrunch == 0 : d.tscrunch % 16 == 0, FBF_ERR_UNSUPPORTED,
                "fbf_ib: tscrunch must divide 16 or be a multiple of it");
    const int Fo = d.nchans / d.fscrunch;
    const long long nwin = T / d.tscrunch;
    const long long nout = nwin * Fo;
    FBF_REQUIRE(tf_bytes == (size_t)nout, FBF_ERR_SIZE,
                "fbf_ib: output size " + std::to_string(tf_bytes) + " != expected " + std::to_string(nout));
    const long long nblocks = (4 * nout + 255) / 256;
    FBF_REQUIRE(nblocks < (1ll << 30) && T / 16 < (1ll << 31), FBF_ERR_UNSUPPORTED, "fbf_ib: grid too large");
    *out = IbFinishJob{partials, static_cast<int8_t*>(tf), plan->d_levels, plan->d_levels + Fo, nwin, (int)(T / 16),
                       Fo, d.tscrunch, d.fscrunch, d.quant_mode, (int)nblocks};
    return FBF_OK;
}

int fbf::ib_fuse_params(const fbf_ib_plan* plan, IbFuseParams* out)
{
    FBF_REQUIRE(plan && out, FBF_ERR_ARG, "fbf_ib: null plan");
    const int fo = plan->d.nchans / plan->d.fscrunch;
    out->offsets = plan->d_levels;
    out->scales = plan->d_levels + fo;
    out->nantennas = plan->d.nantennas;
    out->nchans = plan->d.nchans;
    out->nsamples_per_heap = plan->d.nsamples_per_heap;
    out->tscrunch = plan->d.tscrunch;
    out->fscrunch = plan->d.fscrunch;
    out->quant = plan->d.quant_mode;
    return FBF_OK;
}

extern "C" int fbf_ib_beamform(fbf_ib_plan* plan, const void* taftp, size_t taftp_bytes, void* tf, size_t tf_bytes,
                               fbf_stream_t stream)
{
    return ib_run(plan, IB_INT8, taftp, taftp_bytes, tf, tf_bytes, stream);
}

extern "C" int fbf_ib_beamform_weights(fbf_ib_plan* plan, const void* taftp, size_t taftp_bytes, void* tf,
                                       size_t tf_bytes, fbf_weights_plan* weights_plan, const void* delays,
                                       size_t delays_bytes, double epoch, void* weights, size_t weights_bytes,
                                       fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_ib_beamform_weights: null plan");
    WeightsJob wj{};
    int st = weights_launch_params(weights_plan, delays, delays_bytes, epoch, weights, weights_bytes, &wj.w);
    if (st) return st;
    FBF_REQUIRE(wj.w.F == plan->d.nchans, FBF_ERR_ARG,
                "fbf_ib_beamform_weights: the weights plan's channels differ from the block's");
    FBF_REQUIRE((unsigned long long)wj.w.gx * wj.w.gy < (1ull << 30), FBF_ERR_UNSUPPORTED,
                "fbf_ib_beamform_weights: too many weights work-groups");
    wj.nblocks = (int)(wj.w.gx * wj.w.gy);
    return ib_run(plan, IB_INT8, taftp, taftp_bytes, tf, tf_bytes, stream, wj);
}

extern "C" int fbf_ib_power(fbf_ib_plan* plan, const void* taftp, size_t taftp_bytes, double* power,
                            size_t power_bytes, fbf_stream_t stream)
{
    return ib_run(plan, IB_POWER, taftp, taftp_bytes, power, power_bytes, stream);
}
