// Split transpose TAFTP -> FTPA for a contiguous antenna subset.
//
// Behaviour follows SplitTranspose.cu:11-126 (reference kernel and host
// wrapper): out[f][hg*S + s][p][a] = in[hg][start + a][f][s][p].
//
// MI355X design: for one (heap group, channel) the input is A_used rows of
// 2*S complex samples (1 KiB each for S = 256) at a stride of F KiB, and the
// output is one contiguous [2*S][A_used] block.  A 256-thread block moves a
// 64-antenna x 64-(sample,pol) tile: 16-byte coalesced row loads into LDS,
// then each thread gathers U antennas of one (sample, pol) and writes them as
// one 2*U-byte vector, so the output side is fully coalesced too.
//
// A launch can also generate the block's beamforming weights: its first
// work-groups run the weights generator (weights_run) and leave, and the
// rest do the transpose.  The generator's arithmetic then runs beside the
// transpose's memory traffic, in the same launch, with no dependency
// between the two.
#include "fbf_internal.hpp"

namespace fbf {
namespace {

constexpr int kTileA = 64;    // antennas per LDS tile
constexpr int kTileSP = 64;   // (sample, pol) elements per LDS tile
constexpr int kPad = 2;       // int16 padding per LDS row (row stride 132 B)

template <int U>
struct VecOf;
template <> struct VecOf<8> { using T = int4; };
template <> struct VecOf<4> { using T = int2; };
template <> struct VecOf<2> { using T = int; };
template <> struct VecOf<1> { using T = int16_t; };

// Incoherent beam computed from the same tiles (all antennas of the block):
// output [sample_out][channel] int8, tscrunch dividing the tile's 32 samples,
// fscrunch 1 (IncoherentBeamformer.cu:10-69).
struct IbTile {
    int8_t* out;            // null: no incoherent beam
    const float* offsets;
    const float* scales;
    int tscrunch, quant;
};

template <int U, bool IB>
__global__ __launch_bounds__(256) void split_transpose_k(
    const int16_t* __restrict__ in, int16_t* __restrict__ out,
    int nheap_groups, int total_nantennas, int used_nantennas, int start_antenna,
    int nchans, int sp_per_heap, int n_sp_tiles, IbTile ib, WeightsJob wj)
{
    __shared__ int16_t tile[kTileA][kTileSP + kPad];
    __shared__ int ibsum[4][kTileSP / 2];
    const int tid = threadIdx.x;
    if (weights_job_run(wj, tid)) return;   // uniform over the work-group
    const int logical = xcd_remap((int)blockIdx.x - wj.nblocks, (int)gridDim.x - wj.nblocks);
    const int spt = logical % n_sp_tiles;
    const int rest = logical / n_sp_tiles;
    const int hg = rest % nheap_groups;
    const int f = rest / nheap_groups;
    if (f >= nchans) return;

    const size_t out_block = ((size_t)f * nheap_groups + hg) * sp_per_heap + (size_t)spt * kTileSP;
    // Incoherent beam: a 16-byte row load holds 4 samples x 2 pols of one
    // antenna, and one sdot4 of a 4-byte word with itself is the power of
    // both pols of a sample.  This thread sums its 4 samples over the rows it
    // loads.  A word squares to at most 4 * 128^2 = 2^16, so even the sum of
    // a sample over all A antennas (below) is exact in int32 for A < 2^15,
    // which the incoherent-beam plan requires.
    int ibacc[4] = {0, 0, 0, 0};

    for (int ab = 0; ab < used_nantennas; ab += kTileA) {
        const int na = min(kTileA, used_nantennas - ab);
        // 64 rows x 64 int16 = 8 threads of 16 B per row, 32 rows per pass.
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            const int r = pass * 32 + (tid >> 3);
            const int col = (tid & 7) * 8;
            if (r < na) {
                const size_t row = ((size_t)hg * total_nantennas + start_antenna + ab + r) * nchans + f;
                const int4 v = *reinterpret_cast<const int4*>(in + row * sp_per_heap + (size_t)spt * kTileSP + col);
                if (IB) {
                    ibacc[0] = __builtin_amdgcn_sdot4(v.x, v.x, ibacc[0], false);
                    ibacc[1] = __builtin_amdgcn_sdot4(v.y, v.y, ibacc[1], false);
                    ibacc[2] = __builtin_amdgcn_sdot4(v.z, v.z, ibacc[2], false);
                    ibacc[3] = __builtin_amdgcn_sdot4(v.w, v.w, ibacc[3], false);
                }
                uint32_t* dst = reinterpret_cast<uint32_t*>(&tile[r][col]);
                dst[0] = (uint32_t)v.x;
                dst[1] = (uint32_t)v.y;
                dst[2] = (uint32_t)v.z;
                dst[3] = (uint32_t)v.w;
            }
        }
        __syncthreads();
        const int units = na / U;
        for (int idx = tid; idx < kTileSP * units; idx += 256) {
            const int sp = idx / units;
            const int u = idx - sp * units;
            union {
                typename VecOf<U>::T v;
                int16_t e[U];
            } pack;
#pragma unroll
            for (int i = 0; i < U; ++i) pack.e[i] = tile[u * U + i][sp];
            int16_t* dst = out + (out_block + sp) * used_nantennas + ab + u * U;
            *reinterpret_cast<typename VecOf<U>::T*>(dst) = pack.v;
        }
        __syncthreads();
    }
    if (IB) {
        // Lanes with the same tid % 8 hold the same 4 samples for different
        // rows: sum over the 8 of them in the wave, then over the 4 waves.
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ibacc[q] += __shfl_xor(ibacc[q], 8);
            ibacc[q] += __shfl_xor(ibacc[q], 16);
            ibacc[q] += __shfl_xor(ibacc[q], 32);
        }
        const int lane = tid & 63;
        if (lane < 8) {
#pragma unroll
            for (int q = 0; q < 4; ++q) ibsum[tid >> 6][lane * 4 + q] = ibacc[q];
        }
        __syncthreads();
        // 32 samples per tile -> 32 / tscrunch outputs.
        const int nout = (kTileSP / 2) / ib.tscrunch;
        if (tid < nout) {
            long long P = 0;
            for (int c = tid * ib.tscrunch; c < (tid + 1) * ib.tscrunch; ++c)
                P += (long long)ibsum[0][c] + ibsum[1][c] + ibsum[2][c] + ibsum[3][c];
            const long long t0 = ((long long)hg * sp_per_heap + (long long)spt * kTileSP) / 2;
            const long long to = t0 / ib.tscrunch + tid;
            ib.out[to * nchans + f] = quantise((double)P, ib.offsets[f], ib.scales[f], ib.quant);
        }
    }
}

}  // namespace
}  // namespace fbf

using namespace fbf;

namespace {

template <bool IB>
int st_launch_t(const void* taftp, void* ftpa, int total_nantennas, int used_nantennas, int start_antenna,
                int nchans, int nsamples_per_heap, int nhg, IbTile ib, WeightsJob wj, fbf_stream_t stream)
{
    const int sp_per_heap = 2 * nsamples_per_heap;
    const int n_sp_tiles = sp_per_heap / kTileSP;
    const size_t nblocks = (size_t)nchans * nhg * n_sp_tiles + (size_t)wj.nblocks;
    FBF_REQUIRE(nblocks < (1ull << 31), FBF_ERR_UNSUPPORTED, "fbf_split_transpose: grid too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const auto* in = static_cast<const int16_t*>(taftp);
    auto* out = static_cast<int16_t*>(ftpa);
    dim3 grid((unsigned)nblocks), block(256);
    if (used_nantennas % 8 == 0)
        split_transpose_k<8, IB><<<grid, block, 0, s>>>(in, out, nhg, total_nantennas, used_nantennas,
                                                         start_antenna, nchans, sp_per_heap, n_sp_tiles, ib, wj);
    else if (used_nantennas % 4 == 0)
        split_transpose_k<4, IB><<<grid, block, 0, s>>>(in, out, nhg, total_nantennas, used_nantennas,
                                                         start_antenna, nchans, sp_per_heap, n_sp_tiles, ib, wj);
    else if (used_nantennas % 2 == 0)
        split_transpose_k<2, IB><<<grid, block, 0, s>>>(in, out, nhg, total_nantennas, used_nantennas,
                                                         start_antenna, nchans, sp_per_heap, n_sp_tiles, ib, wj);
    else
        split_transpose_k<1, IB><<<grid, block, 0, s>>>(in, out, nhg, total_nantennas, used_nantennas,
                                                         start_antenna, nchans, sp_per_heap, n_sp_tiles, ib, wj);
    FBF_LAUNCH_CHECK("split_transpose_k");
    return FBF_OK;
}

int st_launch(const void* taftp, void* ftpa, int total_nantennas, int used_nantennas, int start_antenna, int nchans,
              int nsamples_per_heap, int nhg, IbTile ib, fbf_stream_t stream, WeightsJob wj = WeightsJob{})
{
    if (ib.out)
        return st_launch_t<true>(taftp, ftpa, total_nantennas, used_nantennas, start_antenna, nchans,
                                 nsamples_per_heap, nhg, ib, wj, stream);
    return st_launch_t<false>(taftp, ftpa, total_nantennas, used_nantennas, start_antenna, nchans,
                              nsamples_per_heap, nhg, ib, wj, stream);
}

}  // namespace

extern "C" int fbf_split_transpose(const void* taftp, size_t taftp_bytes, void* ftpa,
                                   size_t ftpa_bytes, int total_nantennas, int used_nantennas,
                                   int start_antenna, int nchans, int nsamples_per_heap,
                                   fbf_stream_t stream)
{
    FBF_REQUIRE(taftp && ftpa, FBF_ERR_ARG, "fbf_split_transpose: null buffer");
    FBF_REQUIRE(total_nantennas > 0 && used_nantennas > 0 && start_antenna >= 0 &&
                    start_antenna + used_nantennas <= total_nantennas,
                FBF_ERR_ARG, "fbf_split_transpose: antenna subset outside the input");
    FBF_REQUIRE(nchans > 0 && nsamples_per_heap > 0, FBF_ERR_ARG,
                "fbf_split_transpose: nchans and nsamples_per_heap must be positive");
    FBF_REQUIRE(nsamples_per_heap % 32 == 0, FBF_ERR_UNSUPPORTED,
                "fbf_split_transpose: nsamples_per_heap must be a multiple of 32");
    const size_t hg_bytes = (size_t)4 * total_nantennas * nchans * nsamples_per_heap;
    FBF_REQUIRE(taftp_bytes % hg_bytes == 0, FBF_ERR_SIZE,
                "fbf_split_transpose: input is not a whole number of heap groups");
    const size_t nhg = taftp_bytes / hg_bytes;
    const size_t expect_out = (size_t)4 * used_nantennas * nchans * nsamples_per_heap * nhg;
    FBF_REQUIRE(ftpa_bytes == expect_out, FBF_ERR_SIZE,
                "fbf_split_transpose: output size " + std::to_string(ftpa_bytes) +
                    " != expected " + std::to_string(expect_out));
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(taftp) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(ftpa) & 15) == 0,
                FBF_ERR_ARG, "fbf_split_transpose: buffers must be 16-byte aligned");
    if (nhg == 0) return FBF_OK;
    FBF_REQUIRE(nhg < (1u << 20), FBF_ERR_UNSUPPORTED, "fbf_split_transpose: too many heap groups");

    return st_launch(taftp, ftpa, total_nantennas, used_nantennas, start_antenna, nchans, nsamples_per_heap,
                     (int)nhg, IbTile{nullptr, nullptr, nullptr, 1, 0}, stream);
}

// Checks a fused split-transpose + incoherent-beam call and returns the
// launch arguments.
static int st_ib_params(const void* taftp, size_t taftp_bytes, const void* ftpa, size_t ftpa_bytes, int nchans,
                        int nsamples_per_heap, const fbf_ib_plan* ib_plan, void* tf, size_t tf_bytes, int* nhg_out,
                        int* nantennas, IbTile* tile)
{
    FBF_REQUIRE(taftp && ftpa && ib_plan && tf, FBF_ERR_ARG, "fbf_split_transpose_ib: null argument");
    IbFuseParams ibp;
    int st = ib_fuse_params(ib_plan, &ibp);
    if (st) return st;
    FBF_REQUIRE(ibp.nchans == nchans && ibp.nsamples_per_heap == nsamples_per_heap, FBF_ERR_ARG,
                "fbf_split_transpose_ib: incoherent-beam plan geometry differs from the block's");
    FBF_REQUIRE(ibp.fscrunch == 1 && ibp.tscrunch > 0 && (kTileSP / 2) % ibp.tscrunch == 0, FBF_ERR_UNSUPPORTED,
                "fbf_split_transpose_ib: needs fscrunch 1 and a tscrunch dividing 32");
    const int A = ibp.nantennas;
    FBF_REQUIRE(nsamples_per_heap > 0 && nsamples_per_heap % 32 == 0, FBF_ERR_UNSUPPORTED,
                "fbf_split_transpose_ib: nsamples_per_heap must be a multiple of 32");
    const size_t hg_bytes = (size_t)4 * A * nchans * nsamples_per_heap;
    FBF_REQUIRE(taftp_bytes % hg_bytes == 0 && ftpa_bytes == taftp_bytes, FBF_ERR_SIZE,
                "fbf_split_transpose_ib: input and output must be the same whole number of heap groups");
    const size_t nhg = taftp_bytes / hg_bytes;
    const size_t expect_tf = nhg * nsamples_per_heap / ibp.tscrunch * nchans;
    FBF_REQUIRE(tf_bytes == expect_tf, FBF_ERR_SIZE,
                "fbf_split_transpose_ib: incoherent-beam output size " + std::to_string(tf_bytes) +
                    " != expected " + std::to_string(expect_tf));
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(taftp) & 15) == 0 && (reinterpret_cast<uintptr_t>(ftpa) & 15) == 0,
                FBF_ERR_ARG, "fbf_split_transpose_ib: buffers must be 16-byte aligned");
    FBF_REQUIRE(nhg < (1u << 20), FBF_ERR_UNSUPPORTED, "fbf_split_transpose_ib: too many heap groups");
    *nhg_out = (int)nhg;
    *nantennas = A;
    *tile = IbTile{static_cast<int8_t*>(tf), ibp.offsets, ibp.scales, ibp.tscrunch, ibp.quant};
    return FBF_OK;
}

extern "C" int fbf_split_transpose_ib(const void* taftp, size_t taftp_bytes, void* ftpa, size_t ftpa_bytes,
                                      int nchans, int nsamples_per_heap, fbf_ib_plan* ib_plan, void* tf,
                                      size_t tf_bytes, fbf_stream_t stream)
{
    int nhg = 0, A = 0;
    IbTile ib{};
    const int st = st_ib_params(taftp, taftp_bytes, ftpa, ftpa_bytes, nchans, nsamples_per_heap, ib_plan, tf,
                                tf_bytes, &nhg, &A, &ib);
    if (st) return st;
    if (nhg == 0) return FBF_OK;
    return st_launch(taftp, ftpa, A, A, 0, nchans, nsamples_per_heap, nhg, ib, stream);
}

extern "C" int fbf_split_transpose_ib_weights(const void* taftp, size_t taftp_bytes, void* ftpa, size_t ftpa_bytes,
                                              int nchans, int nsamples_per_heap, fbf_ib_plan* ib_plan, void* tf,
                                              size_t tf_bytes, fbf_weights_plan* weights_plan, const void* delays,
                                              size_t delays_bytes, double epoch, void* weights,
                                              size_t weights_bytes, fbf_stream_t stream)
{
    int nhg = 0, A = 0;
    IbTile ib{};
    int st = st_ib_params(taftp, taftp_bytes, ftpa, ftpa_bytes, nchans, nsamples_per_heap, ib_plan, tf, tf_bytes,
                          &nhg, &A, &ib);
    if (st) return st;
    WeightsJob wj{};
    st = weights_launch_params(weights_plan, delays, delays_bytes, epoch, weights, weights_bytes, &wj.w);
    if (st) return st;
    FBF_REQUIRE(wj.w.F == nchans, FBF_ERR_ARG,
                "fbf_split_transpose_ib_weights: the weights plan's channels differ from the block's");
    FBF_REQUIRE((unsigned long long)wj.w.gx * wj.w.gy < (1ull << 30), FBF_ERR_UNSUPPORTED,
                "fbf_split_transpose_ib_weights: too many weights work-groups");
    wj.nblocks = (int)(wj.w.gx * wj.w.gy);
    if (nhg == 0) return fbf_weights_generate(weights_plan, delays, delays_bytes, epoch, weights, weights_bytes,
                                              nullptr, stream);
    return st_launch(taftp, ftpa, A, A, 0, nchans, nsamples_per_heap, nhg, ib, stream, wj);
}
