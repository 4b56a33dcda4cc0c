// corner_turn.hip -- the byte reorders either side of the beamformer.
//
//   fbf_tuse_transpose  TUSE beam corner turn after it
//                       (meerkat/tuse/src/transpose_to_dada.cpp:19-48)
//   fbf_feng_to_dada    F-engine heap reorder before it
//                       (meerkat/tools/src/feng_to_dada.cu:54-80)
//
// Both are, per batch, the transpose of an R x C matrix of E-byte elements.
// corner_turn_k stages a T x T tile of elements through LDS so that global
// reads are whole input rows (T elements of one row, contiguous) and global
// writes whole output rows, moving 16-, 4- or 1-byte units.  HBM-bound:
// every byte is read once and written once.
#include "fbf_internal.hpp"

#include <climits>

namespace fbf {
namespace {

struct CtArgs {
    const uint8_t* in;
    uint8_t* out;
    long long n2;                  // batch index z = i1 * n2 + i2
    long long in_s1, in_s2;        // byte offsets of batch (i1, i2)
    long long out_s1, out_s2;
    int R, C;                      // input rows and columns (elements)
    int EU;                        // units per element
    int T;                         // tile edge (elements)
    int tiles_r, tiles_c;
    int sh_ru, sh_eu;              // log2(T * EU), log2(EU) (when powers of two)
};

// Tile row pitch in units: T * EU + 1, so that the transposed reads of the
// write phase (threads stepping down the tile rows) fall in distinct banks.
template <typename U, bool POW2>
__global__ __launch_bounds__(256) void corner_turn_k(CtArgs a)
{
    extern __shared__ __align__(16) uint8_t ct_lds[];
    U* tile = reinterpret_cast<U*>(ct_lds);
    const int EU = a.EU;
    const int pitch = a.T * EU + 1;
    long long bid = blockIdx.x;
    const int tc = (int)(bid % a.tiles_c);
    bid /= a.tiles_c;
    const int tr = (int)(bid % a.tiles_r);
    bid /= a.tiles_r;
    const long long i2 = bid % a.n2;
    const long long i1 = bid / a.n2;
    const U* in = reinterpret_cast<const U*>(a.in + i1 * a.in_s1 + i2 * a.in_s2);
    U* out = reinterpret_cast<U*>(a.out + i1 * a.out_s1 + i2 * a.out_s2);
    const int r0 = tr * a.T, c0 = tc * a.T;
    const int nr = min(a.T, a.R - r0), nc = min(a.T, a.C - c0);
    const size_t in_row = (size_t)a.C * EU, out_row = (size_t)a.R * EU;
    const U* src = in + (size_t)r0 * in_row + (size_t)c0 * EU;
    U* dst = out + (size_t)c0 * out_row + (size_t)r0 * EU;

    if (POW2 && nr == a.T && nc == a.T) {
        // Full tile, T * EU and EU powers of two: shifts instead of divisions.
        const int n = a.T << a.sh_ru;
        const int mu = (1 << a.sh_ru) - 1, me = EU - 1;
        for (int i = threadIdx.x; i < n; i += 256) {
            const int r = i >> a.sh_ru, k = i & mu;
            tile[r * pitch + k] = src[(size_t)r * in_row + k];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 256) {
            const int c = i >> a.sh_ru, k = i & mu;
            const int r = k >> a.sh_eu, e = k & me;
            dst[(size_t)c * out_row + k] = tile[r * pitch + c * EU + e];
        }
    } else {
        const int ru = nc * EU;
        for (int i = threadIdx.x; i < nr * ru; i += 256) {
            const int r = i / ru, k = i - r * ru;
            tile[r * pitch + k] = src[(size_t)r * in_row + k];
        }
        __syncthreads();
        const int ou = nr * EU;
        for (int i = threadIdx.x; i < nc * ou; i += 256) {
            const int c = i / ou, k = i - c * ou;
            const int r = k / EU, e = k - r * EU;
            dst[(size_t)c * out_row + k] = tile[r * pitch + c * EU + e];
        }
    }
}

// 4-byte elements, full 64 x 64 tiles: 16-byte global loads (4 words of an
// input row) and stores (4 words of an output row); the tile stays in
// 4-byte LDS words with a 65-word pitch, which keeps both the row writes
// and the column reads free of bank conflicts.  Edge tiles take the general
// path of corner_turn_k.
__global__ __launch_bounds__(256) void corner_turn_w4_k(CtArgs a)
{
    extern __shared__ __align__(16) uint8_t ct_lds[];
    uint32_t* tile = reinterpret_cast<uint32_t*>(ct_lds);
    constexpr int T = 64, P = 65;
    long long bid = blockIdx.x;
    const int tc = (int)(bid % a.tiles_c);
    bid /= a.tiles_c;
    const int tr = (int)(bid % a.tiles_r);
    bid /= a.tiles_r;
    const long long i2 = bid % a.n2;
    const long long i1 = bid / a.n2;
    const uint32_t* in = reinterpret_cast<const uint32_t*>(a.in + i1 * a.in_s1 + i2 * a.in_s2);
    uint32_t* out = reinterpret_cast<uint32_t*>(a.out + i1 * a.out_s1 + i2 * a.out_s2);
    const int r0 = tr * T, c0 = tc * T;
    const int nr = min(T, a.R - r0), nc = min(T, a.C - c0);
    const uint32_t* src = in + (size_t)r0 * a.C + c0;
    uint32_t* dst = out + (size_t)c0 * a.R + r0;
    if (nr == T && nc == T) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = threadIdx.x + 256 * j;
            const int r = i >> 4, c = (i & 15) << 2;
            const uint4 v = *reinterpret_cast<const uint4*>(src + (size_t)r * a.C + c);
            uint32_t* l = tile + r * P + c;
            l[0] = v.x;
            l[1] = v.y;
            l[2] = v.z;
            l[3] = v.w;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = threadIdx.x + 256 * j;
            const int c = i >> 4, r = (i & 15) << 2;
            uint4 v;
            v.x = tile[r * P + c];
            v.y = tile[(r + 1) * P + c];
            v.z = tile[(r + 2) * P + c];
            v.w = tile[(r + 3) * P + c];
            *reinterpret_cast<uint4*>(dst + (size_t)c * a.R + r) = v;
        }
    } else {
        for (int i = threadIdx.x; i < nr * nc; i += 256) {
            const int r = i / nc, c = i - r * nc;
            tile[r * P + c] = src[(size_t)r * a.C + c];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nr * nc; i += 256) {
            const int c = i / nr, r = i - c * nr;
            dst[(size_t)c * a.R + r] = tile[r * P + c];
        }
    }
}

inline int ilog2_exact(long long v)
{
    if (v <= 0 || (v & (v - 1)) != 0) return -1;
    int s = 0;
    while ((1ll << s) < v) ++s;
    return s;
}

// Largest tile edge (power of two, at most 64) whose rows hold at most 64
// units and whose tile fits the LDS budget.
constexpr size_t kCtLdsMax = 32 * 1024;

inline size_t ct_tile_bytes(int T, int EU, size_t U)
{
    return (size_t)T * ((size_t)T * EU + 1) * U;
}

template <typename U>
int ct_launch_u(CtArgs a, long long n1, int E, hipStream_t s)
{
    a.EU = E / (int)sizeof(U);
    int T = 64;
    while (T > 1 && ((long long)T * a.EU > 64 || ct_tile_bytes(T, a.EU, sizeof(U)) > kCtLdsMax)) T >>= 1;
    a.T = T;
    const size_t shm = ct_tile_bytes(T, a.EU, sizeof(U));
    FBF_REQUIRE(shm <= kCtLdsMax, FBF_ERR_UNSUPPORTED, "corner turn: element too large (at most 32 KiB - 16 B)");
    a.tiles_r = (a.R + T - 1) / T;
    a.tiles_c = (a.C + T - 1) / T;
    const long long nblocks = n1 * a.n2 * a.tiles_r * a.tiles_c;
    if (nblocks == 0) return FBF_OK;
    FBF_REQUIRE(nblocks < INT_MAX, FBF_ERR_UNSUPPORTED, "corner turn: too many tiles for one launch");
    a.sh_ru = ilog2_exact((long long)T * a.EU);
    a.sh_eu = ilog2_exact(a.EU);
    const bool pow2 = a.sh_ru >= 0 && a.sh_eu >= 0;
    if (pow2)
        corner_turn_k<U, true><<<dim3((unsigned)nblocks), 256, shm, s>>>(a);
    else
        corner_turn_k<U, false><<<dim3((unsigned)nblocks), 256, shm, s>>>(a);
    FBF_HIP_TRY(hipGetLastError());
    return FBF_OK;
}

// 16-byte units when every element and base is 16-byte aligned, else 4, else 1.
int ct_launch(CtArgs a, long long n1, int E, hipStream_t s)
{
    const uintptr_t al = reinterpret_cast<uintptr_t>(a.in) | reinterpret_cast<uintptr_t>(a.out) |
                         (uintptr_t)a.in_s1 | (uintptr_t)a.in_s2 | (uintptr_t)a.out_s1 | (uintptr_t)a.out_s2;
    if (E == 4 && ((al | (uintptr_t)a.R * 4 | (uintptr_t)a.C * 4) & 15) == 0) {
        // word transposes with 16-byte global accesses (rows 16-byte aligned)
        a.EU = 1;
        a.T = 64;
        a.tiles_r = (a.R + 63) / 64;
        a.tiles_c = (a.C + 63) / 64;
        const long long nblocks = n1 * a.n2 * a.tiles_r * a.tiles_c;
        if (nblocks == 0) return FBF_OK;
        FBF_REQUIRE(nblocks < INT_MAX, FBF_ERR_UNSUPPORTED, "corner turn: too many tiles for one launch");
        corner_turn_w4_k<<<dim3((unsigned)nblocks), 256, 64 * 65 * 4, s>>>(a);
        FBF_HIP_TRY(hipGetLastError());
        return FBF_OK;
    }
    if (E % 16 == 0 && (al & 15) == 0) return ct_launch_u<uint4>(a, n1, E, s);
    if (E % 4 == 0 && (al & 3) == 0) return ct_launch_u<uint32_t>(a, n1, E, s);
    return ct_launch_u<uint8_t>(a, n1, E, s);
}

}  // namespace
}  // namespace fbf

using namespace fbf;

extern "C" int fbf_tuse_transpose(const void* in, size_t in_bytes, void* out, size_t out_bytes, int nchans,
                                  int nsamples, int nfreq, int nbeams, int ngroups, int beam0, int nbeams_out,
                                  fbf_stream_t stream)
{
    FBF_REQUIRE(nchans > 0 && nsamples > 0 && nfreq > 0 && nbeams > 0 && ngroups > 0, FBF_ERR_ARG,
                "fbf_tuse_transpose: geometry values must be positive");
    FBF_REQUIRE(beam0 >= 0 && nbeams_out >= 0 && beam0 + nbeams_out <= nbeams, FBF_ERR_ARG,
                "fbf_tuse_transpose: beam range outside [0, nbeams)");
    const long long per_beam = (long long)nchans * nsamples * nfreq;   // one beam of one group
    const long long need_in = per_beam * nbeams * ngroups;
    const long long need_out = per_beam * ngroups * nbeams_out;
    FBF_REQUIRE((long long)in_bytes >= need_in, FBF_ERR_SIZE,
                "fbf_tuse_transpose: input holds " + std::to_string(in_bytes) + " bytes, the geometry needs " +
                    std::to_string(need_in));
    FBF_REQUIRE((long long)out_bytes == need_out, FBF_ERR_SIZE,
                "fbf_tuse_transpose: output must be nbeams_out * ngroups * nsamples * nfreq * nchans = " +
                    std::to_string(need_out) + " bytes");
    if (nbeams_out == 0) return FBF_OK;
    FBF_REQUIRE(in && out, FBF_ERR_ARG, "fbf_tuse_transpose: null buffer");
    CtArgs a{};
    // batch (beam bo, group g): input [g][beam0 + bo][f][s][c], output [bo][g][s][f][c]
    a.in = static_cast<const uint8_t*>(in) + (size_t)beam0 * per_beam;
    a.out = static_cast<uint8_t*>(out);
    a.n2 = ngroups;
    a.in_s1 = per_beam;
    a.in_s2 = per_beam * nbeams;
    a.out_s1 = per_beam * ngroups;
    a.out_s2 = per_beam;
    a.R = nfreq;
    a.C = nsamples;
    return ct_launch(a, nbeams_out, nchans, static_cast<hipStream_t>(stream));
}

extern "C" int fbf_feng_to_dada(const void* in, void* out, size_t bytes, int nchans, fbf_stream_t stream)
{
    FBF_REQUIRE(nchans > 0, FBF_ERR_ARG, "fbf_feng_to_dada: nchans must be positive");
    const size_t per_ts = (size_t)nchans * 256 * 4;   // MEERKAT_FENG_NSAMPS_PER_HEAP x 2 pols x char2
    FBF_REQUIRE(bytes % per_ts == 0, FBF_ERR_SIZE,
                "fbf_feng_to_dada: number of bytes is not a multiple of the bytes per timestamp "
                "(feng_to_dada.cu:30-35)");
    if (bytes == 0) return FBF_OK;
    FBF_REQUIRE(in && out, FBF_ERR_ARG, "fbf_feng_to_dada: null buffer");
    FBF_REQUIRE(in != out, FBF_ERR_ARG, "fbf_feng_to_dada: input and output must be distinct buffers");
    CtArgs a{};
    a.in = static_cast<const uint8_t*>(in);
    a.out = static_cast<uint8_t*>(out);
    a.n2 = 1;
    a.in_s1 = (long long)per_ts;
    a.out_s1 = (long long)per_ts;
    a.R = nchans;
    a.C = 256;
    return ct_launch(a, (long long)(bytes / per_ts), 4, static_cast<hipStream_t>(stream));
}
