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

template <int U>
__global__ __launch_bounds__(256) void split_transpose_k(
    const int16_t* __restrict__ in, int16_t* __restrict__ out,
    int nheap_groups, int total_nantennas, int used_nantennas, int start_antenna,
    int nchans, int sp_per_heap, int n_sp_tiles)
{
    __shared__ int16_t tile[kTileA][kTileSP + kPad];
    const int tid = threadIdx.x;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int spt = logical % n_sp_tiles;
    const int rest = logical / n_sp_tiles;
    const int hg = rest % nheap_groups;
    const int f = rest / nheap_groups;
    if (f >= nchans) return;

    const size_t out_block = ((size_t)f * nheap_groups + hg) * sp_per_heap + (size_t)spt * kTileSP;

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
}

}  // namespace
}  // namespace fbf

using namespace fbf;

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

    const int sp_per_heap = 2 * nsamples_per_heap;
    const int n_sp_tiles = sp_per_heap / kTileSP;
    const size_t nblocks = (size_t)nchans * nhg * n_sp_tiles;
    FBF_REQUIRE(nblocks < (1ull << 31), FBF_ERR_UNSUPPORTED, "fbf_split_transpose: grid too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const auto* in = static_cast<const int16_t*>(taftp);
    auto* out = static_cast<int16_t*>(ftpa);
    dim3 grid((unsigned)nblocks), block(256);
    if (used_nantennas % 8 == 0)
        split_transpose_k<8><<<grid, block, 0, s>>>(in, out, (int)nhg, total_nantennas, used_nantennas,
                                                     start_antenna, nchans, sp_per_heap, n_sp_tiles);
    else if (used_nantennas % 4 == 0)
        split_transpose_k<4><<<grid, block, 0, s>>>(in, out, (int)nhg, total_nantennas, used_nantennas,
                                                     start_antenna, nchans, sp_per_heap, n_sp_tiles);
    else if (used_nantennas % 2 == 0)
        split_transpose_k<2><<<grid, block, 0, s>>>(in, out, (int)nhg, total_nantennas, used_nantennas,
                                                     start_antenna, nchans, sp_per_heap, n_sp_tiles);
    else
        split_transpose_k<1><<<grid, block, 0, s>>>(in, out, (int)nhg, total_nantennas, used_nantennas,
                                                     start_antenna, nchans, sp_per_heap, n_sp_tiles);
    FBF_LAUNCH_CHECK("split_transpose_k");
    return FBF_OK;
}
