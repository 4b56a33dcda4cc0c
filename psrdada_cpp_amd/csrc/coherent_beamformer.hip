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
// cross-lane traffic.  Power is accumulated exactly in int64 with
// v_mad_i64_i32 (one instruction per squared int32 sum; |r| < 2^22 for
// A <= 128) and handed to the quantiser as a double, like the oracle.
#include "fbf_internal.hpp"

#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

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
    // Voltage columns are `vstride` bytes apart and this beamformer's
    // 2*A bytes start `voff` bytes into each column (vstride = 2*A and
    // voff = 0 for a split-transposed subset; a subarray reading a full
    // FTPA block uses the block's column stride and its antenna offset).
    int vstride, voff;
};

// Loads of one 16-byte K slice [k0, k0 + 16) of a column of K bytes.  They
// are unconditional and branch-free (divergent branches around loads force
// the compiler to wait for every load in flight): addresses are clamped
// inside the column instead.  K is a multiple of 8; AL16 means K % 16 == 0.
//
// Voltage slices: bytes at k >= K may hold anything, because the weights
// there are zero.  A slice past the column end re-reads the column start; the
// upper half of a slice straddling K (K % 16 == 8) re-reads its lower half.
template <bool AL16>
__device__ __forceinline__ v4i load_vslice(const int8_t* col, int k0, int K)
{
    if (AL16) {
        const int kk = k0 < K ? k0 : 0;
        return *reinterpret_cast<const v4i*>(col + kk);
    } else {
        const int klo = k0 < K ? k0 : 0;
        const int khi = k0 + 8 < K ? k0 + 8 : klo;
        const int2 lo = *reinterpret_cast<const int2*>(col + klo);
        const int2 hi = *reinterpret_cast<const int2*>(col + khi);
        return v4i{lo.x, lo.y, hi.x, hi.y};
    }
}

// Weight slices: bytes at k >= K, and every byte of a row past the last
// beam (row_valid false), are zero.
template <bool AL16>
__device__ __forceinline__ v4i load_wslice(const int8_t* col, int k0, int K, bool row_valid)
{
    const v4i r = load_vslice<AL16>(col, k0, K);
    const bool lo_ok = row_valid && k0 < K;
    const bool hi_ok = row_valid && k0 + 8 < K;
    return v4i{lo_ok ? r.x : 0, lo_ok ? r.y : 0, hi_ok ? r.z : 0, hi_ok ? r.w : 0};
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

template <int ROWT, int KSTEPS, int MODE, bool AL16>
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
    const int beam_base = bc * (8 * ROWT);
    const long long t_out = (long long)tc * 16 + c;
    const bool t_valid = t_out < p.nsamples_out;
    // Lanes past the last output sample read (and never store) the last one.
    const long long t_addr = t_valid ? t_out : p.nsamples_out - 1;
    const int tsp = 2 * p.ts;
    const long long TP = 2 * p.nsamples;

    long long pw[ROWT][2];
#pragma unroll
    for (int r = 0; r < ROWT; ++r) pw[r][0] = pw[r][1] = 0;

    for (int fc = 0; fc < p.fs; ++fc) {
        const int f = fo * p.fs + fc;
        // Embedded weights for this channel: lane holds row c of each 16-row tile.
        v4i w[ROWT][KSTEPS];
#pragma unroll
        for (int r = 0; r < ROWT; ++r) {
            const int b = beam_base + r * 8 + (c >> 1);
            const bool bvalid = b < p.B;
            const int8_t* wcol = p.fba + ((size_t)f * p.B + (bvalid ? b : p.B - 1)) * K;
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) {
                const v4i raw = load_wslice<AL16>(wcol, s * 64 + g * 16, K, bvalid);
                v4i e;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    e[q] = (int)((c & 1) ? swap_pairs((uint32_t)raw[q]) : neg_odd((uint32_t)raw[q]));
                w[r][s] = e;
            }
        }

        const int8_t* vcol = p.ftpa + ((size_t)f * TP + (size_t)t_addr * tsp) * p.vstride + p.voff;
        v4i v[KSTEPS];
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) v[s] = load_vslice<AL16>(vcol, s * 64 + g * 16, K);

        for (int j = 0; j < tsp; ++j) {
            // Prefetch the next column while this one is multiplied (the last
            // step re-reads its own column).
            v4i vn[KSTEPS];
            const int jn = j + 1 < tsp ? j + 1 : j;
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) vn[s] = load_vslice<AL16>(vcol + (size_t)jn * p.vstride, s * 64 + g * 16, K);
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
                pw[r][0] += (long long)acc[0] * acc[0];
                pw[r][0] += (long long)acc[1] * acc[1];
                pw[r][1] += (long long)acc[2] * acc[2];
                pw[r][1] += (long long)acc[3] * acc[3];
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
                dst[b0] = (double)pw[r][0];
                dst[b0 + 1] = (double)pw[r][1];
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
            out[heap_base + (size_t)b0 * tf_size] = quantise((double)pw[r][0], off, sc, p.quant);
            out[heap_base + (size_t)(b0 + 1) * tf_size] = quantise((double)pw[r][1], off, sc, p.quant);
        }
    }
}

// ---------------------------------------------------------------------------
// v2: v_mfma_i32_32x32x32_i8 (the full-rate int8 MFMA on gfx950).
//
// A (32 rows x 32 k) = embedded weights, lane l: row l%32, K half h = l/32.
// B (32 k x 32 cols) = voltages, lane l: column l%32, K half h.
// D lane l: column l%32, rows 8*(i/4) + 4*h + i%4 for i = 0..15, so a lane
// holds Re and Im of beams 4*q + 2*h + {0, 1} (q = 0..3) of its row tile.
// K is split into 16-byte slices; slice (h, s) covers bytes
// [16 * (h * KSTEPS + s), +16) of a column, so each lane reads one contiguous
// 16*KSTEPS-byte half of every voltage column.  The 32 columns of a tile are
// 32 output samples; column j of the integration window is processed at
// step j, with two columns in flight (va / vb) so that loads overlap the
// MFMAs and the FP64 detection of the previous column.
// ---------------------------------------------------------------------------
typedef int v16i __attribute__((ext_vector_type(16)));

template <int KSTEPS, bool AL16>
__device__ __forceinline__ void cb32_load_col(v4i (&v)[KSTEPS], const int8_t* col, int h, int K)
{
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) v[s] = load_vslice<AL16>(col, 16 * (h * KSTEPS + s), K);
}

template <int RT, int KSTEPS, int MODE>
__device__ __forceinline__ void cb32_column(const v4i (&w)[RT][KSTEPS], const v4i (&v)[KSTEPS],
                                            long long (&pw)[RT][4][2], const CbArgs& p, int f, long long n,
                                            int beam_base, int h, bool t_valid)
{
    v16i acc[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        acc[r] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s)
            acc[r] = __builtin_amdgcn_mfma_i32_32x32x32_i8(w[r][s], v[s], acc[r], 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        if (MODE == MODE_SUMS) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
                if (t_valid && b0 < p.B) {
                    int32_t* dst = static_cast<int32_t*>(p.out) +
                                   (((size_t)f * 2 * p.nsamples + (size_t)n) * p.B + b0) * 2;
                    *reinterpret_cast<v4i*>(dst) = v4i{acc[r][4 * q], acc[r][4 * q + 1], acc[r][4 * q + 2],
                                                       acc[r][4 * q + 3]};
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int xr = acc[r][4 * q + 2 * e];
                const int xi = acc[r][4 * q + 2 * e + 1];
                pw[r][q][e] += (long long)xr * xr;
                pw[r][q][e] += (long long)xi * xi;
            }
        }
    }
}

// MFMA half of cb32_column: acc[r] = W'[r] x column.
template <int RT, int KSTEPS>
__device__ __forceinline__ void cb32_mfma(v16i (&acc)[RT], const v4i (&w)[RT][KSTEPS], const v4i (&v)[KSTEPS])
{
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        acc[r] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s)
            acc[r] = __builtin_amdgcn_mfma_i32_32x32x32_i8(w[r][s], v[s], acc[r], 0, 0, 0);
    }
}

// Detection half of cb32_column: |beam|^2 into the exact int64 accumulators
// (and, in MODE_SUMS, the raw int32 sums to memory).
template <int RT, int MODE>
__device__ __forceinline__ void cb32_detect(const v16i (&acc)[RT], long long (&pw)[RT][4][2], const CbArgs& p,
                                            int f, long long n, int beam_base, int h, bool t_valid)
{
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        if (MODE == MODE_SUMS) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
                if (t_valid && b0 < p.B) {
                    int32_t* dst = static_cast<int32_t*>(p.out) +
                                   (((size_t)f * 2 * p.nsamples + (size_t)n) * p.B + b0) * 2;
                    *reinterpret_cast<v4i*>(dst) = v4i{acc[r][4 * q], acc[r][4 * q + 1], acc[r][4 * q + 2],
                                                       acc[r][4 * q + 3]};
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int xr = acc[r][4 * q + 2 * e];
                const int xi = acc[r][4 * q + 2 * e + 1];
                pw[r][q][e] += (long long)xr * xr;
                pw[r][q][e] += (long long)xi * xi;
            }
    }
}

template <int KSTEPS>
__device__ __forceinline__ void lds_col(v4i (&v)[KSTEPS], const int8_t* src, int jj, int c, int h, int kp)
{
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s)
        v[s] = *reinterpret_cast<const v4i*>(src + (jj * 32 + c) * kp + 16 * (h * KSTEPS + s));
}

template <int RT, int KSTEPS, int MODE, bool AL16>
__global__ __launch_bounds__(256) void cb_mfma32_k(CbArgs p)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;
    const int c = lane & 31;
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
    const int beam_base = bc * (16 * RT);
    const long long t_out = (long long)tc * 32 + c;
    const bool t_valid = t_out < p.nsamples_out;
    // Lanes past the last output sample read (and never store) the last one.
    const long long t_addr = t_valid ? t_out : p.nsamples_out - 1;
    const int tsp = 2 * p.ts;
    const long long TP = 2 * p.nsamples;

    long long pw[RT][4][2];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) pw[r][q][0] = pw[r][q][1] = 0;

    for (int fc = 0; fc < p.fs; ++fc) {
        const int f = fo * p.fs + fc;
        v4i w[RT][KSTEPS];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int b = beam_base + r * 16 + (c >> 1);
            const bool bvalid = b < p.B;
            const int8_t* wcol = p.fba + ((size_t)f * p.B + (bvalid ? b : p.B - 1)) * K;
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) {
                const v4i raw = load_wslice<AL16>(wcol, 16 * (h * KSTEPS + s), K, bvalid);
                v4i e;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    e[q] = (int)((c & 1) ? swap_pairs((uint32_t)raw[q]) : neg_odd((uint32_t)raw[q]));
                w[r][s] = e;
            }
        }

        const int8_t* vcol = p.ftpa + ((size_t)f * TP + (size_t)t_addr * tsp) * p.vstride + p.voff;
        const long long n0 = t_out * tsp;
        v4i va[KSTEPS], vb[KSTEPS];
        cb32_load_col<KSTEPS, AL16>(va, vcol, h, K);
        cb32_load_col<KSTEPS, AL16>(vb, vcol + p.vstride, h, K);
        for (int j = 0; j < tsp; j += 2) {
            // Two columns in flight; the last steps re-read the final column.
            const int j2 = j + 2 < tsp ? j + 2 : tsp - 1;
            const int j3 = j + 3 < tsp ? j + 3 : tsp - 1;
            v4i cur[KSTEPS];
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) cur[s] = va[s];
            cb32_load_col<KSTEPS, AL16>(va, vcol + (size_t)j2 * p.vstride, h, K);
            cb32_column<RT, KSTEPS, MODE>(w, cur, pw, p, f, n0 + j, beam_base, h, t_valid);
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) cur[s] = vb[s];
            cb32_load_col<KSTEPS, AL16>(vb, vcol + (size_t)j3 * p.vstride, h, K);
            cb32_column<RT, KSTEPS, MODE>(w, cur, pw, p, f, n0 + j + 1, beam_base, h, t_valid);
        }
    }

    if (MODE == MODE_SUMS || !t_valid) return;
    if (MODE == MODE_POWER) {
        double* dst = static_cast<double*>(p.out) + ((size_t)fo * p.nsamples_out + t_out) * p.B;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
                if (b0 < p.B) {
                    dst[b0] = (double)pw[r][q][0];
                    dst[b0 + 1] = (double)pw[r][q][1];
                }
            }
        return;
    }
    const float off = p.offsets[fo];
    const float sc = p.scales[fo];
    int8_t* out = static_cast<int8_t*>(p.out);
    const size_t tf_size = (size_t)p.ohs * p.Fo;
    const size_t heap_base = (size_t)(t_out / p.ohs) * p.B * tf_size + (size_t)(t_out % p.ohs) * p.Fo + fo;
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
            if (b0 < p.B) {
                out[heap_base + (size_t)b0 * tf_size] = quantise((double)pw[r][q][0], off, sc, p.quant);
                out[heap_base + (size_t)(b0 + 1) * tf_size] = quantise((double)pw[r][q][1], off, sc, p.quant);
            }
        }
}

// ---------------------------------------------------------------------------
// v3: 32x32x32 MFMA with the voltage tile staged through LDS.
//
// The four waves of a work-group share the same 32 output samples and take
// 16*RT beams each, so one copy of the voltage columns in LDS feeds all four.
// The integration window of the 32 samples is streamed in chunks of JC
// columns: every thread loads its share of the next chunk from HBM/L2 with
// wide coalesced loads (each sample's JC columns are contiguous) while the
// current chunk is multiplied out of the other LDS buffer.  LDS rows are
// padded to 32*KSTEPS + 16 bytes so that the ds_read_b128 of consecutive
// lanes fall on distinct bank groups.
// ---------------------------------------------------------------------------
constexpr int kLdsMaxPieces = 4;   // per-thread prefetch registers (pieces)

template <int RT, int KSTEPS, int MODE, bool AL16>
__global__ __launch_bounds__(256) void cb_lds_k(CbArgs p, int jc)
{
    extern __shared__ __align__(16) int8_t lds[];
    constexpr int KP = 32 * KSTEPS + 16;
    constexpr int PB = AL16 ? 16 : 8;   // bytes per fill piece
    using Piece = typename std::conditional<AL16, v4i, int2>::type;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int h = lane >> 5;
    const int c = lane & 31;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int bgroup = logical % p.n_bgroups;
    const int rest = logical / p.n_bgroups;
    const int tgroup = rest % p.n_tgroups;
    const int fo = rest / p.n_tgroups;
    if (fo >= p.Fo) return;   // uniform over the work-group

    const int K = 2 * p.A;
    const int beam_base = (bgroup * 4 + wave) * (16 * RT);
    const bool wave_active = beam_base < p.B;
    const long long t0 = (long long)tgroup * 32;
    const long long t_out = t0 + c;
    const bool t_valid = t_out < p.nsamples_out;
    const int tsp = 2 * p.ts;
    const long long TP = 2 * p.nsamples;
    const int buf_bytes = jc * 32 * KP;
    const int pieces_per_run = jc * K / PB;   // one output sample's jc columns
    const int npieces = 32 * pieces_per_run;
    const int nchunks = tsp / jc;

    long long pw[RT][4][2];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) pw[r][q][0] = pw[r][q][1] = 0;

    for (int fc = 0; fc < p.fs; ++fc) {
        const int f = fo * p.fs + fc;
        v4i w[RT][KSTEPS];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int b = beam_base + r * 16 + (c >> 1);
            const bool bvalid = b < p.B;
            const int8_t* wcol = p.fba + ((size_t)f * p.B + (bvalid ? b : p.B - 1)) * K;
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) {
                const v4i raw = load_wslice<AL16>(wcol, 16 * (h * KSTEPS + s), K, bvalid);
                v4i e;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    e[q] = (int)((c & 1) ? swap_pairs((uint32_t)raw[q]) : neg_odd((uint32_t)raw[q]));
                w[r][s] = e;
            }
        }

        const int8_t* vbase = p.ftpa + (size_t)f * TP * p.vstride + p.voff;
        Piece pre[kLdsMaxPieces];
        auto fetch = [&](int chunk) {
#pragma unroll
            for (int u = 0; u < kLdsMaxPieces; ++u) {
                const int i = tid + 256 * u;
                const int ii = i < npieces ? i : npieces - 1;
                const int cc = ii / pieces_per_run;
                const int o = (ii - cc * pieces_per_run) * PB;
                long long tc = t0 + cc;
                tc = tc < p.nsamples_out ? tc : p.nsamples_out - 1;
                const int jj = o / K;
                pre[u] = *reinterpret_cast<const Piece*>(vbase + ((size_t)tc * tsp + (size_t)chunk * jc + jj) * p.vstride +
                                                         (o - jj * K));
            }
        };
        auto stash = [&](int buf) {
            int8_t* dst = lds + buf * buf_bytes;
#pragma unroll
            for (int u = 0; u < kLdsMaxPieces; ++u) {
                const int i = tid + 256 * u;
                if (i < npieces) {
                    const int cc = i / pieces_per_run;
                    const int o = (i - cc * pieces_per_run) * PB;
                    const int jj = o / K;
                    const int kb = o - jj * K;
                    *reinterpret_cast<Piece*>(dst + (jj * 32 + cc) * KP + kb) = pre[u];
                }
            }
        };

        fetch(0);
        stash(0);
        __syncthreads();
        for (int k = 0; k < nchunks; ++k) {
            if (k + 1 < nchunks) fetch(k + 1);
            if (wave_active) {
                // One column at a time: with three waves per SIMD the
                // hardware interleaves MFMA and detection of different waves
                // (a two-deep software pipeline costs a wave of occupancy and
                // measured slower).
                const int8_t* src = lds + (k & 1) * buf_bytes;
                const long long n0 = t_out * tsp + (long long)k * jc;
                for (int jj = 0; jj < jc; ++jj) {
                    v4i v[KSTEPS];
                    lds_col<KSTEPS>(v, src, jj, c, h, KP);
                    v16i acc[RT];
                    cb32_mfma<RT, KSTEPS>(acc, w, v);
                    cb32_detect<RT, MODE>(acc, pw, p, f, n0 + jj, beam_base, h, t_valid);
                }
            }
            if (k + 1 < nchunks) stash((k + 1) & 1);
            __syncthreads();
        }
    }

    if (MODE == MODE_SUMS || !t_valid || !wave_active) return;
    if (MODE == MODE_POWER) {
        double* dst = static_cast<double*>(p.out) + ((size_t)fo * p.nsamples_out + t_out) * p.B;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
                if (b0 < p.B) {
                    dst[b0] = (double)pw[r][q][0];
                    dst[b0 + 1] = (double)pw[r][q][1];
                }
            }
        return;
    }
    const float off = p.offsets[fo];
    const float sc = p.scales[fo];
    int8_t* out = static_cast<int8_t*>(p.out);
    const size_t tf_size = (size_t)p.ohs * p.Fo;
    const size_t heap_base = (size_t)(t_out / p.ohs) * p.B * tf_size + (size_t)(t_out % p.ohs) * p.Fo + fo;
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
            if (b0 < p.B) {
                out[heap_base + (size_t)b0 * tf_size] = quantise((double)pw[r][q][0], off, sc, p.quant);
                out[heap_base + (size_t)(b0 + 1) * tf_size] = quantise((double)pw[r][q][1], off, sc, p.quant);
            }
        }
}

// Chunk length: the largest power of two <= 16 that divides tsp and keeps
// one chunk within kLdsMaxPieces pieces per thread.
inline int cb_lds_chunk(int tsp, int K, bool al16)
{
    const int pb = al16 ? 16 : 8;
    int jc = 16;
    while (jc > 1 && (tsp % jc != 0 || 32 * jc * K / pb > 256 * kLdsMaxPieces)) jc >>= 1;
    return jc;
}

inline size_t cb_lds_bytes(int ksteps, int jc) { return (size_t)2 * jc * 32 * (32 * ksteps + 16); }

template <int RT, int MODE, bool AL16>
hipError_t launch_lds_rt(const CbArgs& a, int ksteps, int jc, dim3 grid, hipStream_t s)
{
    const size_t shm = cb_lds_bytes(ksteps, jc);
    switch (ksteps) {
    case 1: cb_lds_k<RT, 1, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 2: cb_lds_k<RT, 2, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 3: cb_lds_k<RT, 3, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 4: cb_lds_k<RT, 4, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 5: cb_lds_k<RT, 5, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 6: cb_lds_k<RT, 6, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 7: cb_lds_k<RT, 7, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    case 8: cb_lds_k<RT, 8, MODE, AL16><<<grid, 256, shm, s>>>(a, jc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_lds_mode(const CbArgs& a, int rt, int ksteps, int jc, dim3 grid, hipStream_t s)
{
    const bool al16 = (a.A % 8) == 0;
    switch (rt) {
    case 1: return al16 ? launch_lds_rt<1, MODE, true>(a, ksteps, jc, grid, s)
                        : launch_lds_rt<1, MODE, false>(a, ksteps, jc, grid, s);
    case 2: return al16 ? launch_lds_rt<2, MODE, true>(a, ksteps, jc, grid, s)
                        : launch_lds_rt<2, MODE, false>(a, ksteps, jc, grid, s);
    case 4: return al16 ? launch_lds_rt<4, MODE, true>(a, ksteps, jc, grid, s)
                        : launch_lds_rt<4, MODE, false>(a, ksteps, jc, grid, s);
    default: return hipErrorInvalidValue;
    }
}

template <int RT, int MODE, bool AL16>
hipError_t launch32_rt(const CbArgs& a, int ksteps, dim3 grid, hipStream_t s)
{
    switch (ksteps) {
    case 1: cb_mfma32_k<RT, 1, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 2: cb_mfma32_k<RT, 2, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 3: cb_mfma32_k<RT, 3, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 4: cb_mfma32_k<RT, 4, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 5: cb_mfma32_k<RT, 5, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 6: cb_mfma32_k<RT, 6, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 7: cb_mfma32_k<RT, 7, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 8: cb_mfma32_k<RT, 8, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch32_mode(const CbArgs& a, int rt, int ksteps, dim3 grid, hipStream_t s)
{
    const bool al16 = (a.A % 8) == 0;
    if (rt == 4)
        return al16 ? launch32_rt<4, MODE, true>(a, ksteps, grid, s) : launch32_rt<4, MODE, false>(a, ksteps, grid, s);
    return al16 ? launch32_rt<2, MODE, true>(a, ksteps, grid, s) : launch32_rt<2, MODE, false>(a, ksteps, grid, s);
}

template <int ROWT, int MODE, bool AL16>
hipError_t launch_rowt(const CbArgs& a, int ksteps, dim3 grid, hipStream_t s)
{
    switch (ksteps) {
    case 1: cb_mfma_k<ROWT, 1, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 2: cb_mfma_k<ROWT, 2, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 3: cb_mfma_k<ROWT, 3, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    case 4: cb_mfma_k<ROWT, 4, MODE, AL16><<<grid, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const CbArgs& a, int rowt, int ksteps, dim3 grid, hipStream_t s)
{
    const bool al16 = (a.A % 8) == 0;
    if (rowt == 8)
        return al16 ? launch_rowt<8, MODE, true>(a, ksteps, grid, s) : launch_rowt<8, MODE, false>(a, ksteps, grid, s);
    return al16 ? launch_rowt<4, MODE, true>(a, ksteps, grid, s) : launch_rowt<4, MODE, false>(a, ksteps, grid, s);
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
    int chunk;        // LDS chunk length in columns, 0 = automatic
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

// Process-wide autotuning table: geometry -> (variant, chunk).
namespace {
std::mutex g_tune_mutex;
std::map<std::string, std::pair<int, int>> g_tune_table;

std::string tune_key(const fbf_cb_desc& d)
{
    return std::to_string(d.nantennas) + "/" + std::to_string(d.nbeams) + "/" + std::to_string(d.nchans) + "/" +
           std::to_string(d.tscrunch) + "/" + std::to_string(d.fscrunch);
}
}  // namespace

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
    auto* p = new fbf_cb_plan{*desc, ohs, 0, 0, nullptr};
    hipError_t e = hipMalloc(&p->d_levels, sizeof(float) * 2 * Fo);
    if (e != hipSuccess) {
        delete p;
        return hip_status(e, "hipMalloc(levels)");
    }
    st = fbf_cb_plan_set_levels(p, offsets, scales);
    if (st) {
        (void)hipFree(p->d_levels);
        delete p;
        return st;
    }
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_table.find(tune_key(*desc));
        if (it != g_tune_table.end()) {
            p->variant = it->second.first;
            p->chunk = it->second.second;
        }
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

// Kernel variants: MFMA shape and beams per wave.  Variant 0 picks
// automatically (the tuned choice for the geometry).
struct CbVariant {
    int mfma;        // 32: v_mfma_i32_32x32x32_i8, 16: v_mfma_i32_16x16x64_i8
    int rows;        // row tiles per wave (RT for 32x32, ROWT for 16x16)
    bool lds;        // voltage tile staged through LDS and shared by the work-group
};
static const CbVariant kCbVariants[] = {
    {0, 0, false},   // auto
    {32, 4, false},  // 64 beams x 32 samples per wave
    {32, 2, false},  // 32 beams x 32 samples per wave
    {16, 8, false},  // 64 beams x 16 samples per wave
    {16, 4, false},  // 32 beams x 16 samples per wave
    {32, 1, true},   // LDS: 16 beams per wave, 64 per work-group
    {32, 2, true},   // LDS: 32 beams per wave, 128 per work-group
    {32, 4, true},   // LDS: 64 beams per wave, 256 per work-group
};

static CbVariant cb_pick_variant(const fbf_cb_desc& d, int variant)
{
    if (variant > 0) return kCbVariants[variant];
    (void)d;
    return kCbVariants[5];
}

extern "C" int fbf_cb_variant_count(int* count)
{
    FBF_REQUIRE(count, FBF_ERR_ARG, "fbf_cb_variant_count: null pointer");
    *count = (int)(sizeof(kCbVariants) / sizeof(kCbVariants[0]));
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

// Validates and launches one coherent-beam kernel over T samples of voltage
// columns that are `vstride` bytes apart, the beamformer's antennas starting
// `voff` bytes into each column.
static int cb_launch(fbf_cb_plan* plan, int mode, const void* ftpa, long long T, int vstride, int voff,
                     const void* fba, size_t weights_bytes, void* out, size_t out_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb: null plan");
    FBF_REQUIRE(ftpa && fba && out, FBF_ERR_ARG, "fbf_cb: null buffer");
    const fbf_cb_desc& d = plan->d;
    FBF_REQUIRE(T % d.tscrunch == 0, FBF_ERR_SIZE, "fbf_cb: nsamples is not a multiple of tscrunch");
    FBF_REQUIRE(vstride >= 2 * d.nantennas + voff && voff >= 0 && vstride % 8 == 0 && voff % 8 == 0,
                FBF_ERR_ARG, "fbf_cb: bad voltage column stride/offset");
    const bool al16_view = (vstride % 16 == 0) && (voff % 16 == 0);
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

    const CbVariant var = cb_pick_variant(d, plan->variant);
    const int kbytes = var.mfma == 32 ? 32 : 64;
    const int ksteps = (2 * d.nantennas + kbytes - 1) / kbytes;
    const int beams_per_wave = var.mfma == 32 ? 16 * var.rows : 8 * var.rows;
    const int t_per_wave = var.mfma == 32 ? 32 : 16;
    const bool al16 = d.nantennas % 8 == 0;
    FBF_REQUIRE(!al16 || al16_view, FBF_ERR_ARG,
                "fbf_cb: a multiple-of-8 antenna subset needs a 16-byte aligned column stride and offset");
    int jc = cb_lds_chunk(2 * d.tscrunch, 2 * d.nantennas, al16);
    if (plan->chunk > 0) {
        FBF_REQUIRE((2 * d.tscrunch) % plan->chunk == 0 && plan->chunk <= jc, FBF_ERR_ARG,
                    "fbf_cb: chunk must divide 2*tscrunch and fit the prefetch registers");
        jc = plan->chunk;
    }
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
    a.vstride = vstride;
    a.voff = voff;
    a.n_bchunks = (d.nbeams + beams_per_wave - 1) / beams_per_wave;
    a.n_tchunks = (int)((To + t_per_wave - 1) / t_per_wave);
    if (var.lds) {
        // Four waves share the samples and split the beams.
        a.wb = 4;
        a.n_bgroups = (a.n_bchunks + 3) / 4;
        a.n_tgroups = a.n_tchunks;
    } else {
        a.wb = a.n_bchunks >= 4 ? 4 : (a.n_bchunks >= 2 ? 2 : 1);
        a.n_bgroups = (a.n_bchunks + a.wb - 1) / a.wb;
        const int wt = 4 / a.wb;
        a.n_tgroups = (a.n_tchunks + wt - 1) / wt;
    }
    const long long nblocks = (long long)Fo * a.n_tgroups * a.n_bgroups;
    FBF_REQUIRE(nblocks > 0 && nblocks < (1ll << 31), FBF_ERR_UNSUPPORTED, "fbf_cb: grid too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    dim3 grid((unsigned)nblocks);
    hipError_t e;
    if (var.lds) {
        if (mode == MODE_INT8)
            e = launch_lds_mode<MODE_INT8>(a, var.rows, ksteps, jc, grid, s);
        else if (mode == MODE_SUMS)
            e = launch_lds_mode<MODE_SUMS>(a, var.rows, ksteps, jc, grid, s);
        else
            e = launch_lds_mode<MODE_POWER>(a, var.rows, ksteps, jc, grid, s);
    } else if (var.mfma == 32) {
        if (mode == MODE_INT8)
            e = launch32_mode<MODE_INT8>(a, var.rows, ksteps, grid, s);
        else if (mode == MODE_SUMS)
            e = launch32_mode<MODE_SUMS>(a, var.rows, ksteps, grid, s);
        else
            e = launch32_mode<MODE_POWER>(a, var.rows, ksteps, grid, s);
    } else {
        if (mode == MODE_INT8)
            e = launch_mode<MODE_INT8>(a, var.rows, ksteps, grid, s);
        else if (mode == MODE_SUMS)
            e = launch_mode<MODE_SUMS>(a, var.rows, ksteps, grid, s);
        else
            e = launch_mode<MODE_POWER>(a, var.rows, ksteps, grid, s);
    }
    if (e != hipSuccess) return hip_status(e, "cb_mfma_k launch");
    return FBF_OK;
}

static int cb_run(fbf_cb_plan* plan, int mode, const void* ftpa, size_t ftpa_bytes, const void* fba,
                  size_t weights_bytes, void* out, size_t out_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb: null plan");
    long long T = 0;
    int st = cb_nsamples(plan, ftpa_bytes, &T);
    if (st) return st;
    return cb_launch(plan, mode, ftpa, T, 2 * plan->d.nantennas, 0, fba, weights_bytes, out, out_bytes, stream);
}

extern "C" int fbf_cb_beamform_subarray(fbf_cb_plan* plan, const void* ftpa_block, size_t block_bytes,
                                        int block_nantennas, int antenna_offset, const void* fba_weights,
                                        size_t weights_bytes, void* tbtf, size_t tbtf_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb_beamform_subarray: null plan");
    FBF_REQUIRE(block_nantennas > 0 && antenna_offset >= 0 &&
                    antenna_offset + plan->d.nantennas <= block_nantennas,
                FBF_ERR_ARG, "fbf_cb_beamform_subarray: subarray outside the block's antennas");
    const size_t per_sample = (size_t)2 * 2 * block_nantennas * plan->d.nchans;
    FBF_REQUIRE(block_bytes % per_sample == 0, FBF_ERR_SIZE,
                "fbf_cb_beamform_subarray: block is not a whole number of samples");
    const long long T = (long long)(block_bytes / per_sample);
    return cb_launch(plan, MODE_INT8, ftpa_block, T, 2 * block_nantennas, 2 * antenna_offset, fba_weights,
                     weights_bytes, tbtf, tbtf_bytes, stream);
}

extern "C" int fbf_cb_plan_set_chunk(fbf_cb_plan* plan, int chunk)
{
    FBF_REQUIRE(plan && chunk >= 0 && chunk <= 16 && (chunk & (chunk - 1)) == 0, FBF_ERR_ARG,
                "fbf_cb_plan_set_chunk: chunk must be 0 (auto) or a power of two <= 16");
    plan->chunk = chunk;
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

// ---------------------------------------------------------------------------
// Autotuner: times every (variant, chunk) candidate on synthetic data of the
// plan's geometry.  The voltage values do not change the instruction stream,
// so constant bytes are enough for timing.
// ---------------------------------------------------------------------------
extern "C" int fbf_cb_autotune(fbf_cb_plan* plan, long long nsamples, int reps, int* best_variant, int* best_chunk,
                               float* best_ms)
{
    FBF_REQUIRE(plan && nsamples > 0 && reps > 0, FBF_ERR_ARG, "fbf_cb_autotune: bad argument");
    const fbf_cb_desc& d = plan->d;
    FBF_REQUIRE(nsamples % d.tscrunch == 0 && (nsamples / d.tscrunch) % plan->ohs == 0, FBF_ERR_ARG,
                "fbf_cb_autotune: nsamples must fill whole output heaps");
    const size_t vbytes = (size_t)nsamples * 2 * 2 * d.nantennas * d.nchans;
    const size_t wbytes = (size_t)2 * d.nchans * d.nbeams * d.nantennas;
    const size_t obytes = (size_t)(nsamples / d.tscrunch) * d.nbeams * (d.nchans / d.fscrunch);
    void *v = nullptr, *w = nullptr, *o = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (v) (void)hipFree(v);
        if (w) (void)hipFree(w);
        if (o) (void)hipFree(o);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    };
    hipError_t e = hipMalloc(&v, vbytes);
    if (e == hipSuccess) e = hipMalloc(&w, wbytes);
    if (e == hipSuccess) e = hipMalloc(&o, obytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(v, 0x11, vbytes, s);
    if (e == hipSuccess) e = hipMemsetAsync(w, 0x5A, wbytes, s);
    if (e != hipSuccess) {
        cleanup();
        return hip_status(e, "fbf_cb_autotune setup");
    }
    const int nvar = (int)(sizeof(kCbVariants) / sizeof(kCbVariants[0]));
    const bool al16 = d.nantennas % 8 == 0;
    const int jc_max = cb_lds_chunk(2 * d.tscrunch, 2 * d.nantennas, al16);
    const int saved_variant = plan->variant, saved_chunk = plan->chunk;
    int bv = saved_variant, bc = saved_chunk;
    float bms = 1e30f;
    for (int var = 1; var < nvar; ++var) {
        std::vector<int> chunks{0};
        if (kCbVariants[var].lds)
            for (int jc = jc_max / 2; jc >= 1; jc >>= 1) chunks.push_back(jc);
        for (int jc : chunks) {
            plan->variant = var;
            plan->chunk = jc;
            int st = cb_launch(plan, MODE_INT8, v, nsamples, 2 * d.nantennas, 0, w, wbytes, o, obytes, s);
            if (st != FBF_OK) continue;   // candidate not applicable to this geometry
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps; ++r)
                cb_launch(plan, MODE_INT8, v, nsamples, 2 * d.nantennas, 0, w, wbytes, o, obytes, s);
            (void)hipEventRecord(e1, s);
            e = hipEventSynchronize(e1);
            if (e != hipSuccess) {
                plan->variant = saved_variant;
                plan->chunk = saved_chunk;
                cleanup();
                return hip_status(e, "fbf_cb_autotune timing");
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)reps;
            if (ms < bms) {
                bms = ms;
                bv = var;
                bc = jc;
            }
        }
    }
    plan->variant = bv;
    plan->chunk = bc;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_table[tune_key(d)] = {bv, bc};
    }
    cleanup();
    if (best_variant) *best_variant = bv;
    if (best_chunk) *best_chunk = bc;
    if (best_ms) *best_ms = bms;
    return FBF_OK;
}
