// Coherent beamformer on CDNA4 int8 MFMA.
//
// Behaviour follows CoherentBeamformer.cu:48-225 (kernel bf_aptf_general_k and
// CoherentBeamformer::beamform):
//   beam(b, f, t, p) = sum_a w[f][b][a] * v[f][t][p][a]      (no conjugate)
//   P(b, fo, to)     = sum over fscrunch, tscrunch, pol of |beam|^2
//   out              = quantise(P, offset[fo], scale[fo])   written TBTF
//
// The int8 output leaves the beamformer kernel as [Fo][To][B] (beams
// contiguous: each work-group writes whole 64-byte runs) and a second kernel
// (tbtf_k) moves it to TBTF [To/ohs][B][ohs][Fo].  Writing TBTF directly from
// the beamformer scatters single bytes 1 channel apart across work-groups
// (one work-group owns one channel), which the memory system turns into
// 32-byte partial writes: 32x the output size in HBM writes.
//
// MI355X design.  Per channel the contraction is a real GEMM
//   C[2B x T*P] = W'[2B x 2A] * V'[2A x T*P]
// with the [Re, -Im; Im, Re] embedding: row Re_b = [wr_a, -wi_a] and
// row Im_b = [wi_a, wr_a] over k = (antenna, re/im).  k is interleaved exactly
// as FTPA stores a column (char2 per antenna), so every 16-byte K slice of a
// voltage column is one contiguous 16-byte load straight into the MFMA B
// operand -- no de-interleave and no LDS staging.  The weights are embedded in
// registers from the raw FBA char2 bytes (byte swap for Im rows, negation
// of the odd bytes for Re rows), once per channel per wave; waves whose
// weights hold a wi = -128, whose negation int8 cannot hold, take an exact
// path instead (cb_exact_outputs), so the sums are exact for every int8
// weight.
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

#include <climits>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace fbf {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

// Two adjacent beams' int8 outputs as one 16-bit store (b is even, B % 8 == 0).
__device__ __forceinline__ void store_pair(int8_t* dst, int8_t a, int8_t b)
{
    *reinterpret_cast<uint16_t*>(dst) = (uint16_t)((uint8_t)a | ((uint16_t)(uint8_t)b << 8));
}

enum { MODE_INT8 = 0, MODE_SUMS = 1, MODE_POWER = 2 };

struct CbArgs {
    const int8_t* ftpa;
    const int8_t* fba;
    void* out;
    const float* offsets;
    const float* scales;
    const long long* thr;   // quantiser thresholds [Fo][kThr] (cb_thresholds_k)
    long long nsamples;   // T
    long long nsamples_out;
    int A, B, F, Fo, ts, fs, ohs, quant;
    int n_bchunks, n_tchunks, n_bgroups, n_tgroups, wb;
    // Voltage columns are `vstride` bytes apart and this beamformer's
    // 2*A bytes start `voff` bytes into each column (vstride = 2*A and
    // voff = 0 for a split-transposed subset; a subarray reading a full
    // FTPA block uses the block's column stride and its antenna offset).
    int vstride, voff;
    // S > 0: the voltages are instead a TAFTP block, [heap group][Atot][F][S
    // samples][2 pols] char2, and this beamformer's antennas start at its
    // antenna a0 (read so by cb_tr_k<..., TA = true> and cb_exact_outputs).
    int S, Atot, a0;
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

// Im row: [wr, wi] -> [wi, wr] per antenna (swap the bytes of each 16-bit
// half: one byte permutation).
__device__ __forceinline__ uint32_t swap_pairs(uint32_t x)
{
    return __builtin_amdgcn_perm(x, x, 0x02030001u);
}

typedef unsigned short us2_t __attribute__((ext_vector_type(2)));

// Re row: [wr, wi] -> [wr, -wi] per antenna.  The odd byte of each 16-bit
// half is negated by a packed 16-bit subtract from zero: its low byte is
// zero, so no borrow reaches the even byte.  -wi is exact for wi != -128
// (wi = -128 gives -128 again); weights holding -128 never reach the
// embedded form (below).
__device__ __forceinline__ uint32_t neg_odd(uint32_t x)
{
    const us2_t n = us2_t{0, 0} - __builtin_bit_cast(us2_t, x & 0xFF00FF00u);
    return (x & 0x00FF00FFu) | __builtin_bit_cast(uint32_t, n);
}

// Embedded 16-byte weight slice of an Im row (swap_pairs) or a Re row
// (neg_odd).  The callers pass im_row as a constant, so only one form is
// compiled.
__device__ __forceinline__ v4i embed_slice(v4i raw, bool im_row)
{
    v4i e;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t x = (uint32_t)raw[q];
        e[q] = (int)(im_row ? swap_pairs(x) : neg_odd(x));
    }
    return e;
}

// Weight components wi = -128.  The Re row would carry -wi = 128, which
// int8 cannot hold, while the reference's int32 sums
// (r = xx - yy, CoherentBeamformer.cu:150-155) are exact for every int8
// weight.  Weights round(127 e^{i theta}) never hold -128, so the MFMA paths
// stay as they are and every kernel first checks the weights it will use:
// a wave (or a work-group, where its waves share staging) that finds a
// wi = -128 computes its outputs with cb_exact_outputs instead.
//
// Flag dword from a dword x: 0x01 at the odd byte of every 16-bit pair of x
// whose even byte is 0x80, else 0.  e's even byte is zero exactly for 0x80,
// and adding 0xFF to it carries into bit 8 exactly when it is not zero (at
// most 0x1FE, so the pairs do not interact).
__device__ __forceinline__ int m128_corr(uint32_t x)
{
    const uint32_t e = (x & 0x00FF00FFu) ^ 0x00800080u;
    return (int)(((e + 0x00FF00FFu) & 0x01000100u) ^ 0x01000100u);
}

// Whether a raw FBA slice [wr, wi, ...] holds a wi = -128 (an odd byte 0x80).
__device__ __forceinline__ bool raw_has_m128(v4i raw)
{
    int any = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) any |= m128_corr((uint32_t)raw[q] >> 8);
    return any != 0;
}

// The same over a wave: uniform, so the exact path is a scalar branch.
__device__ __forceinline__ bool wave_any(bool x)
{
    return __builtin_amdgcn_ballot_w64(x) != 0;
}

// Exact outputs of output channel fo for beams [b0, b0 + nb) and output
// windows [w0, w1): the int32 sums of every column (MODE_SUMS,
// [F][T][P][B][2]), the power (MODE_POWER, double [Fo][To][B]) or the staged
// int8 output (MODE_INT8, [Fo][To][B]).  Thread t of nt takes (beam, window)
// pairs t, t + nt, ... and forms plain int32 dot products:
// r = sum wr vr - wi vi and i = sum wr vi + wi vr are exact for every int8
// weight (CoherentBeamformer.cu:150-155); the power is exact in int64 and the
// quantiser is the exact one.  The path of waves (or work-groups) whose
// weights hold a wi = -128; consecutive threads take consecutive beams, so a
// wave's voltage reads are one broadcast column.
template <int MODE>
__device__ __forceinline__ void cb_exact_outputs(const CbArgs& p, int fo, int b0, int nb, long long w0,
                                                 long long w1, int t, int nt)
{
    if (w1 > p.nsamples_out) w1 = p.nsamples_out;
    if (w1 <= w0) return;
    const int K = 2 * p.A;
    const size_t TP = (size_t)2 * p.nsamples;
    const int W = 2 * p.ts;
    const float off = p.offsets[fo], sc = p.scales[fo];
    const long long n = (long long)nb * (w1 - w0);
    for (long long idx = t; idx < n; idx += nt) {
        const int b = b0 + (int)(idx % nb);
        const long long w = w0 + idx / nb;
        if (b >= p.B) continue;
        long long P = 0;
        for (int fc = 0; fc < p.fs; ++fc) {
            const int f = fo * p.fs + fc;
            const int8_t* wrow = p.fba + ((size_t)f * p.B + b) * K;
            for (int j = 0; j < W; ++j) {
                // Column (sample, pol) jc of channel f; antenna a at col + a * astep.
                const long long jc = w * W + j;
                const int8_t* col;
                size_t astep;
                if (p.S > 0) {
                    const long long t = jc >> 1;
                    col = p.ftpa + (((size_t)(t / p.S) * p.Atot + p.a0) * p.F + f) * p.S * 4 +
                          (size_t)(t % p.S) * 4 + 2 * (jc & 1);
                    astep = (size_t)p.F * p.S * 4;
                } else {
                    col = p.ftpa + ((size_t)f * TP + (size_t)jc) * p.vstride + p.voff;
                    astep = 2;
                }
                int re = 0, im = 0;
                for (int a = 0; a < p.A; ++a) {
                    const int vr = col[a * astep], vi = col[a * astep + 1];
                    const int wr = wrow[2 * a], wi = wrow[2 * a + 1];
                    re += wr * vr - wi * vi;
                    im += wr * vi + wi * vr;
                }
                if (MODE == MODE_SUMS) {
                    int32_t* dst = static_cast<int32_t*>(p.out) + (((size_t)f * TP + (size_t)w * W + j) * p.B + b) * 2;
                    dst[0] = re;
                    dst[1] = im;
                }
                P += (long long)re * re + (long long)im * im;
            }
        }
        const size_t o = ((size_t)fo * p.nsamples_out + (size_t)w) * p.B + b;
        if (MODE == MODE_POWER) static_cast<double*>(p.out)[o] = (double)P;
        if (MODE == MODE_INT8) static_cast<int8_t*>(p.out)[o] = quantise((double)P, off, sc, p.quant);
    }
}

// Whether any weight of beams [b0, b0 + nb) over the fscrunch channels of
// output channel fo holds a wi = -128, as seen by this lane (lanes take
// 16-byte pieces t, t + nt, ...).  K = 2A is a multiple of 8: pieces are
// 16 bytes, or 8 at the end of a row.
__device__ __forceinline__ bool cb_weights_have_m128(const CbArgs& p, int fo, int b0, int nb, int t, int nt)
{
    const int K = 2 * p.A;
    const int kp = (K + 15) / 16;
    bool m = false;
    const int n = p.fs * nb * kp;
    for (int i = t; i < n; i += nt) {
        const int fc = i / (nb * kp);
        const int r = i - fc * nb * kp;
        const int b = b0 + r / kp;
        const int k0 = 16 * (r % kp);
        if (b >= p.B) continue;
        const int8_t* row = p.fba + ((size_t)(fo * p.fs + fc) * p.B + b) * K;
        const int2 lo = *reinterpret_cast<const int2*>(row + k0);
        const int2 hi = k0 + 8 < K ? *reinterpret_cast<const int2*>(row + k0 + 8) : int2{0, 0};
        m |= raw_has_m128(v4i{lo.x, lo.y, hi.x, hi.y});
    }
    return m;
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

    // A wave whose weights hold a wi = -128 takes the exact path.
    if (wave_any(cb_weights_have_m128(p, fo, beam_base, 8 * ROWT, lane, 64))) {
        cb_exact_outputs<MODE>(p, fo, beam_base, 8 * ROWT, (long long)tc * 16, (long long)tc * 16 + 16, lane, 64);
        return;
    }

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
            for (int s = 0; s < KSTEPS; ++s) w[r][s] = load_wslice<AL16>(wcol, s * 64 + g * 16, K, bvalid);
        }
        // Every load is in flight before the first is used.
#pragma unroll
        for (int r = 0; r < ROWT; ++r)
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) w[r][s] = embed_slice(w[r][s], c & 1);

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
    // int8, staged [fo][to][b]
    const float off = p.offsets[fo];
    const float sc = p.scales[fo];
    int8_t* row = static_cast<int8_t*>(p.out) + ((size_t)fo * p.nsamples_out + t_out) * p.B;
#pragma unroll
    for (int r = 0; r < ROWT; ++r) {
        const int b0 = beam_base + r * 8 + 2 * g;
        if (b0 < p.B)
            store_pair(row + b0, quantise((double)pw[r][0], off, sc, p.quant),
                       quantise((double)pw[r][1], off, sc, p.quant));
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

    // A wave whose weights hold a wi = -128 takes the exact path.
    if (wave_any(cb_weights_have_m128(p, fo, beam_base, 16 * RT, lane, 64))) {
        cb_exact_outputs<MODE>(p, fo, beam_base, 16 * RT, (long long)tc * 32, (long long)tc * 32 + 32, lane, 64);
        return;
    }

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
            for (int s = 0; s < KSTEPS; ++s) w[r][s] = load_wslice<AL16>(wcol, 16 * (h * KSTEPS + s), K, bvalid);
        }
        // Every load is in flight before the first is used.
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) w[r][s] = embed_slice(w[r][s], c & 1);

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
    // int8, staged [fo][to][b]
    const float off = p.offsets[fo];
    const float sc = p.scales[fo];
    int8_t* row = static_cast<int8_t*>(p.out) + ((size_t)fo * p.nsamples_out + t_out) * p.B;
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
            if (b0 < p.B)
                store_pair(row + b0, quantise((double)pw[r][q][0], off, sc, p.quant),
                           quantise((double)pw[r][q][1], off, sc, p.quant));
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

    // A work-group any of whose weights hold a wi = -128 takes the exact
    // path (its waves share the voltage staging, so all of them).
    if (__syncthreads_or(cb_weights_have_m128(p, fo, bgroup * 4 * 16 * RT, 4 * 16 * RT, tid, 256) ? 1 : 0)) {
        cb_exact_outputs<MODE>(p, fo, bgroup * 4 * 16 * RT, 4 * 16 * RT, t0, t0 + 32, tid, 256);
        return;
    }

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
            for (int s = 0; s < KSTEPS; ++s) w[r][s] = load_wslice<AL16>(wcol, 16 * (h * KSTEPS + s), K, bvalid);
        }
        // Every load is in flight before the first is used.
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) w[r][s] = embed_slice(w[r][s], c & 1);

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
    // int8, staged [fo][to][b]
    const float off = p.offsets[fo];
    const float sc = p.scales[fo];
    int8_t* row = static_cast<int8_t*>(p.out) + ((size_t)fo * p.nsamples_out + t_out) * p.B;
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b0 = beam_base + r * 16 + 4 * q + 2 * h;
            if (b0 < p.B)
                store_pair(row + b0, quantise((double)pw[r][q][0], off, sc, p.quant),
                           quantise((double)pw[r][q][1], off, sc, p.quant));
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

// ---------------------------------------------------------------------------
// Transposed tiles: voltages as the A operand, weights as the B operand.
//
//   D[m][n] = sum_k V[m][k] W'[n][k]
//
// Rows m are 32 consecutive input columns (t, p) of one channel, columns n
// are 16 beams x (Re, Im) embedding rows.  Lane l holds column n = l % 32
// and rows 8*(i/4) + 4*(l/32) + i%4 (i = 0..15).  For an integration window
// of W = 2*tscrunch = 8, 16 or 32 columns, or a multiple of 32, every window
// a lane touches is a fixed group of its 16 values, so the squares sum into
// one int64 per (beam tile, window) in the lane.  The four lanes holding
// {Re, Im} x {row halves} of a beam are combined once per pass of windows by
// a reduce-scatter (exchange with lane ^ 1, then lane ^ 32) that leaves each
// lane a distinct (beam, window) total to quantise.
//
// A wave keeps RT beam tiles (16 RT beams) of weights in registers for its
// channel and reads each voltage tile once from LDS for all of them: LDS
// traffic is 1/RT of a tile per MFMA.  With fscrunch FS = 2 or 4 the waves of
// a work-group take different channels of the fscrunch group (each with its
// own weights) and add their window totals through LDS before quantising.
//
// Voltage staging is a ring of NB tile sets in LDS (one tile per channel of
// the fscrunch group per set) filled by direct global->LDS loads issued
// NB - 1 tiles ahead.  Every step issues one set (past the work-group's last
// tile the loads repeat the last tile into a slot nobody reads), so the
// number of loads in flight after a wave's loads of the current set is the
// constant (NB - 2) * M and the wait is a fixed s_waitcnt.
// ---------------------------------------------------------------------------

// s_waitcnt vmcnt(N) lgkmcnt(0).
template <int N>
__device__ __forceinline__ void wait_vm_lgkm0()
{
    static_assert(N >= 0 && N <= 63, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4));
}

// One 16-byte global->LDS load per lane: lane i's 16 bytes land at
// lds_run + 16 * i (lds_run is wave-uniform).
__device__ __forceinline__ void dma16(const int8_t* src, int8_t* lds_run)
{
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_run, 16, 0, 0);
}

// A wave-uniform pointer as scalar registers, so that a load from it plus a
// 32-bit lane offset takes the scalar-base address form (no 64-bit vector
// address arithmetic per load).
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* ptr)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(ptr);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}

// A pointer known to address global memory, so that stores through it are
// global (not flat: a flat store also counts against the LDS wait counter).
__device__ __forceinline__ __attribute__((address_space(1))) int8_t* global_ptr(int8_t* ptr)
{
    return (__attribute__((address_space(1))) int8_t*)ptr;
}

// Work-group barrier without the memory fence of __syncthreads (which would
// wait for every outstanding global load, including the LDS prefetches).
// Callers wait for their own LDS and global->LDS traffic first; the empty
// asm statements keep the compiler from moving memory accesses across.
__device__ __forceinline__ void block_barrier()
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

struct CbTrArgs {
    CbArgs p;
    int ntiles;                // 32-column tiles per channel
    int tps;                   // tiles per window (W >= 32), else 1
    int npasses;               // passes per work-group
    int ntp;                   // tiles per pass
    int n_bgroups, n_tgroups;
    // TAFTP staging: bytes per heap group of the block, log2 of the tiles
    // (16 samples) per heap.
    size_t hstride;
    int lgtph;
    // Incoherent-beam partials (cb_tr_k with IBF): [F][ntiles][2 halves][32
    // columns] int32; sample j of a tile has the power over the block's
    // antennas and both pols split over columns 2j, 2j + 1 (its pols) of
    // both halves (ib_finish_run adds the four).
    int* ibp;
};

// The incoherent-beam partials of tiles [t0, t1) of channel f straight from
// the TAFTP block, for the work-groups on the exact path (which leave the
// tile loop): per sample the power over all antennas and both pols in
// column 2j of half 0, zeros in its other three places.
__device__ __forceinline__ void cb_exact_ib_partials(const CbArgs& p, int* ibp, int ntiles, int f, int t0, int t1,
                                                  int tid, int nt)
{
    const size_t astep = (size_t)p.F * p.S * 4;
    for (int i = tid; i < (t1 - t0) * 16; i += nt) {
        const int tile = t0 + i / 16, j = i % 16;
        const long long t = (long long)tile * 16 + j;
        const int8_t* v = p.ftpa + (((size_t)(t / p.S) * p.Atot + p.a0) * p.F + f) * p.S * 4 + (size_t)(t % p.S) * 4;
        int sum = 0;
        for (int a = 0; a < p.A; ++a) {
            const int w = *reinterpret_cast<const int*>(v + a * astep);
            sum = __builtin_amdgcn_sdot4(w, w, sum, false);
        }
        int* dst = ibp + ((size_t)f * ntiles + tile) * 64 + 2 * j;
        dst[0] = sum;
        dst[1] = 0;
        dst[32] = 0;
        dst[33] = 0;
    }
}

// The 16-byte A operands of one step's K slices (8 antennas x char2 of one
// column each) from an LDS tile laid out [antenna][32 columns] int16, 64
// bytes a row: two ds_read_b64_tr_b16 of 4 antenna rows each per slice.
// Within every 16-lane group G the instruction gives lane i, as its element
// r, element i % 4 of the 8 bytes read by lane 16 G + 4 r + i / 4.  So when
// lane L reads the 4 columns 16 (G % 2) + 4 (L % 4) .. + 3 of antenna row
// base + (L / 4) % 4, lane c + 32 h (c < 32) receives column c of rows
// base .. base + 3, in row order.  `src` is that row piece of slice 0 for
// this lane; slice s is 8 rows (512 bytes) further on, its second read 4
// rows (256 bytes) after its first.
//
// All the reads are issued in one asm block that ends with their LDS wait.
// Issued through the builtin, every transposed read is preceded by a wait
// for all outstanding global->LDS loads (the compiler cannot tell that they
// fill other ring slots), which cost the kernel a memory round trip per
// step; the ring's own vmcnt wait and barrier already order them.  The
// outputs are complete when the block ends (lgkmcnt(0) inside it).
template <int KSTEPS>
__device__ __forceinline__ void tr_operands(const int8_t* src, v4i (&a)[KSTEPS])
{
    typedef int v2i __attribute__((ext_vector_type(2)));
    const uint32_t addr = (uint32_t)(size_t)(const __attribute__((address_space(3))) int8_t*)src;
    v2i r[2 * KSTEPS];
    static_assert(KSTEPS >= 1 && KSTEPS <= 4, "K steps");
    if constexpr (KSTEPS == 1) {
        __asm__ volatile(
            "ds_read_b64_tr_b16 %0, %2\n\t"
            "ds_read_b64_tr_b16 %1, %2 offset:256\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(r[0]), "=&v"(r[1])
            : "v"(addr)
            : "memory");
    } else if constexpr (KSTEPS == 2) {
        __asm__ volatile(
            "ds_read_b64_tr_b16 %0, %4\n\t"
            "ds_read_b64_tr_b16 %1, %4 offset:256\n\t"
            "ds_read_b64_tr_b16 %2, %4 offset:512\n\t"
            "ds_read_b64_tr_b16 %3, %4 offset:768\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3])
            : "v"(addr)
            : "memory");
    } else if constexpr (KSTEPS == 3) {
        __asm__ volatile(
            "ds_read_b64_tr_b16 %0, %6\n\t"
            "ds_read_b64_tr_b16 %1, %6 offset:256\n\t"
            "ds_read_b64_tr_b16 %2, %6 offset:512\n\t"
            "ds_read_b64_tr_b16 %3, %6 offset:768\n\t"
            "ds_read_b64_tr_b16 %4, %6 offset:1024\n\t"
            "ds_read_b64_tr_b16 %5, %6 offset:1280\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5])
            : "v"(addr)
            : "memory");
    } else {
        __asm__ volatile(
            "ds_read_b64_tr_b16 %0, %8\n\t"
            "ds_read_b64_tr_b16 %1, %8 offset:256\n\t"
            "ds_read_b64_tr_b16 %2, %8 offset:512\n\t"
            "ds_read_b64_tr_b16 %3, %8 offset:768\n\t"
            "ds_read_b64_tr_b16 %4, %8 offset:1024\n\t"
            "ds_read_b64_tr_b16 %5, %8 offset:1280\n\t"
            "ds_read_b64_tr_b16 %6, %8 offset:1536\n\t"
            "ds_read_b64_tr_b16 %7, %8 offset:1792\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]),
              "=&v"(r[7])
            : "v"(addr)
            : "memory");
    }
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) a[s] = v4i{r[2 * s].x, r[2 * s].y, r[2 * s + 1].x, r[2 * s + 1].y};
}

// A tile's incoherent-beam partials from the operands tr_operands gave this
// wave (every K slice): per lane the power of its column over the 8 * KSTEPS
// antennas of its row half, each word (two antennas' re, im) squared with
// v_dot4, antenna groups of 8 past A left out (they re-read antenna 0); the
// 64 lanes store their sums as the tile's [2 halves][32 columns] record.
template <int KSTEPS>
__device__ __forceinline__ void ib_tile(const v4i (&a)[KSTEPS], int* ibp, int ntiles, int A, int f, int tile,
                                        int lane)
{
    const int h = lane >> 5;
    int v = 0;
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) {
        int t = __builtin_amdgcn_sdot4(a[s].x, a[s].x, 0, false);
        t = __builtin_amdgcn_sdot4(a[s].y, a[s].y, t, false);
        t = __builtin_amdgcn_sdot4(a[s].z, a[s].z, t, false);
        t = __builtin_amdgcn_sdot4(a[s].w, a[s].w, t, false);
        v += 8 * (h * KSTEPS + s) < A ? t : 0;
    }
    if (tile < ntiles) ibp[((size_t)f * ntiles + tile) * 64 + lane] = v;
}

// a + b' where lanes below 32 keep a and take b' = (a of lane + 32), and
// lanes from 32 keep b and take (b of lane - 32): the halves' exchange of a
// reduce-scatter in one v_permlane32_swap per dword.  With a = value for the
// lower half and b = value for the upper half, every lane gets its own sum.
__device__ __forceinline__ long long swap32_sum(long long a, long long b)
{
    const unsigned alo = (uint32_t)(unsigned long long)a, ahi = (uint32_t)((unsigned long long)a >> 32);
    const unsigned blo = (uint32_t)(unsigned long long)b, bhi = (uint32_t)((unsigned long long)b >> 32);
    const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    // 64-bit sum as one add and one add-with-carry of the halves.
    const unsigned slo = lo[0] + lo[1];
    const unsigned shi = hi[0] + hi[1] + (slo < lo[0] ? 1u : 0u);
    return (long long)(((unsigned long long)shi << 32) | slo);
}

// Constants of the threshold quantiser's estimates: FP64 (P - offset) * inv
// with inv = RN(1 / scale), and FP32 P * inv_f + c_f with inv_f = RN(1 /
// scale), c_f = RN(-offset * inv_f).
struct QuantConst {
    double off_d, inv;
    float inv_f, c_f;
};

__device__ __forceinline__ QuantConst quant_const(float off, float sc)
{
    const float inv_f = 1.0f / sc;
    return QuantConst{(double)off, 1.0 / (double)sc, inv_f, -off * inv_f};
}

// Quantiser thresholds.  For a positive, normal, finite scale and a finite
// offset, quantise((double)P, offset, scale, mode) is a non-decreasing step
// function of the integer power P: thr[j] is the least P >= 0 that quantises
// to at least j - 128 (thr[0] = INT64_MIN, thr[256] = INT64_MAX, and
// INT64_MAX where no P < 2^62 gets there).  The kernel needs only an
// estimate e within one step of the result q: the two thresholds around e
// then settle q without a division or a branch (P >= thr[e + 129] means
// q = e + 1, P < thr[e + 128] means q = e - 1).  The estimate is
// e = trunc(clamp(y, -128, 127)) for an approximation y of the quotient Y
// that the exact quantiser rounds: whenever |y - Y| < 0.5, truncation moves
// y by less than 1 and the mode's rounding moves Y by at most 0.5 (truncation
// by less than 1, and then |y - Y| < 1 suffices), so |e - q| < 2 in either
// mode (the clamp keeps this), i.e. e is q or a neighbour.
constexpr int kThr = 257;
constexpr int kThrBytes = 2304;   // kThr int64, padded to a multiple of 256 bytes
static_assert(kThrBytes >= kThr * 8, "threshold table");

__global__ __launch_bounds__(256) void cb_thresholds_k(const float* offsets, const float* scales, int mode,
                                                       long long* thr)
{
    const int fo = blockIdx.x;
    const float off = offsets[fo], sc = scales[fo];
    long long* t = thr + (size_t)fo * kThr;
    const int j = threadIdx.x + 1;   // 1 .. 256
    if (j == 256) {
        t[0] = LLONG_MIN;
        t[256] = LLONG_MAX;
        return;
    }
    const int k = j - 128;
    long long lo = 0, hi = 1ll << 62;
    if (quantise((double)hi, off, sc, mode) < k) {
        t[j] = LLONG_MAX;
        return;
    }
    while (lo < hi) {   // least P in [lo, hi] with quantise(P) >= k
        const long long mid = lo + (hi - lo) / 2;
        if (quantise((double)mid, off, sc, mode) >= k)
            hi = mid;
        else
            lo = mid + 1;
    }
    t[j] = lo;
}

__device__ __forceinline__ bool thr_applies(float off, float sc)
{
    return isfinite(off) && sc >= 1.17549435e-38f && sc <= 3.40282347e+38f;
}

// Threshold correction of an estimate e in [-128, 127]; thr_mid = thr + 128
// (so thr_mid[e] and thr_mid[e + 1] are one ds_read2_b64).
__device__ __forceinline__ int8_t thr_settle(long long P, int e, const long long* thr_mid)
{
    const long long lo = thr_mid[e], hi = thr_mid[e + 1];
    return (int8_t)(e + (int)(P >= hi) - (int)(P < lo));
}

// FP64 estimate RN(RN((double)P - offset) * RN(1 / scale)): it rounds the
// same difference as the exact quantiser, so it is within 2^-43 of Y for
// |Y| <= 256 at any offset and scale; beyond, both saturate alike.
__device__ __forceinline__ int thr_estimate(long long P, const QuantConst& q)
{
    const double y = ((double)P - q.off_d) * q.inv;
    return (int)fmin(fmax(y, -128.0), 127.0);
}

__device__ __forceinline__ int8_t quantise_thr(long long P, const QuantConst& q, const long long* thr_mid)
{
    return thr_settle(P, thr_estimate(P, q), thr_mid);
}

// FP32 estimate (about a third of the VALU time of the FP64 one).  For
// P >= 0, RN(RN(hi) * 2^32 + RN(lo)) is within 3 * 2^-24 of P relative;
// RN(1 / scale) and c = RN(-offset * inv) are within 2^-24 and 2^-23
// relative, and the fma rounds once, so for |Y| <= 130 the estimate is
// within 2^-24 * (4 |P / scale| + 2 |offset / scale| + 131) <=
// 2^-24 * (651 + 6 |offset / scale|) of Y: below 0.38 when
// |offset / scale| < 2^20 (thr_f32_applies); beyond |Y| = 130 both
// saturate alike.
__device__ __forceinline__ bool thr_f32_applies(float off, float sc)
{
    return fabsf(off / sc) < 1048576.0f;
}

__device__ __forceinline__ int thr_estimate_f32(long long P, const QuantConst& q)
{
    const float pf = (float)(int)(P >> 32) * 4294967296.0f + (float)(uint32_t)P;
    const float y = __builtin_fmaf(pf, q.inv_f, q.c_f);
    return (int)__builtin_amdgcn_fmed3f(y, -128.0f, 127.0f);
}

// One tile step of the transposed kernels: the RT tiles' MFMA chains and the
// detection of their values into the pass accumulators pw[RT / 2][NSLOT]
// (tile r adds to beam group r / 2; values of window q of the tile go to
// slot g * SPT + q), one v_mad_i64_i32 per value.  (Splitting the
// dependent chains into independent partial sums measured 0.7 % slower:
// the other waves of the SIMD already cover their latency.)
template <int RT, int KSTEPS, int SPT, int NSLOT>
__device__ __forceinline__ void detect_step(const v4i (&a)[KSTEPS], const v4i (&w)[RT][KSTEPS],
                                            long long (&pw)[RT / 2][NSLOT], int g)
{
    constexpr int n = 16 / SPT;   // values per window per tile
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        v16i acc = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], w[r][s], acc, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < SPT; ++q)
#pragma unroll
            for (int i = 0; i < n; ++i) pw[r >> 1][g * SPT + q] += (long long)acc[q * n + i] * acc[q * n + i];
    }
}

// Staging constants of cb_tr_k for FS channels x KSTEPS 32-byte K steps.
template <int FS, int KSTEPS, int NTS>
struct TrStage {
    static constexpr int NW = 4 * NTS;                     // waves per work-group
    static constexpr int NI = NTS * FS * KSTEPS;           // 1 KiB loads per set
    static constexpr int M = (NI + NW - 1) / NW;           // per wave (the rest load into a dummy run)
    static constexpr int SET = NI * 1024;                  // bytes per set
    static constexpr int NB = 60 / NI < 2 ? 2 : (60 / NI > 8 ? 8 : 60 / NI);
    static constexpr int G = NB >= 4 ? 2 : 1;             // steps per barrier
    static constexpr int WAIT = (NB - 2 * G) * M;
    static constexpr int DUMMY = (NI % NW) ? NW * 1024 : 0;
};

template <int RT, int NSLOT>
struct TrRed {
    // totals per lane after the reduce-scatter
    static constexpr int VALS = RT == 2 ? NSLOT / 2 : (RT / 4) * NSLOT;
};

// RT: beam tiles per wave (RT / 2 groups of 32 beams, a Re and an Im tile
// each); SPT: windows per tile (1, 2, 4); NSLOT: windows
// per pass held in registers; FS: fscrunch; NTS: time slots per work-group;
// TA: the voltages are a TAFTP block (CbArgs::S > 0), staged straight into
// [antenna][column] tiles and read with transposed LDS loads (tr_operands),
// else FTPA columns; IBF (with TA, FS = 1, all antennas of the block): the
// work-groups of beam group 0 also leave the incoherent beam's per-sample
// partials (CbTrArgs::ibp) from the same operands.
//
// A work-group has 4 * NTS waves: NTS time slots (contiguous runs of passes
// of the same channel group and beams) x 4 waves ((4 / FS) beam groups x FS
// channels).  With 64 beams per wave a work-group holds every wave of its
// compute unit (NTS = tr_nts waves per SIMD: 3 or 4), so the per-step
// barrier keeps them in lockstep: waves of separate work-groups sharing a
// SIMD progress at different rates (the issue arbiter prefers older waves),
// which leaves SIMDs with one wave at the end.  With 128 beams per wave
// (two waves per SIMD) a work-group is one time slot, two share a compute
// unit, and they take turns at a raised wave priority instead.
template <int RT, int KSTEPS, int SPT, int NSLOT, int FS, int NTS, bool TA, bool IBF>
__global__ __launch_bounds__(256 * NTS) void cb_tr_k(CbTrArgs ta)
{
    using St = TrStage<FS, KSTEPS, NTS>;
    constexpr int KPAD = 32 * KSTEPS;
    constexpr int PR = 2 * KSTEPS;                 // 16-byte pieces per LDS column
    constexpr int VALS = TrRed<RT, NSLOT>::VALS;
    constexpr int NB = St::NB;
    constexpr int NG = RT / 2;                     // groups of 32 beams per wave
    static_assert(RT == 2 || RT == 4 || RT == 8, "RT");
    static_assert(NSLOT % SPT == 0 && (RT != 2 || NSLOT % 2 == 0), "slots");
    static_assert(St::M <= 2, "at most two staging loads per wave per set");
    static_assert(!IBF || (TA && FS == 1 && KSTEPS <= 4 && NTS <= 2), "incoherent-beam partials");
    extern __shared__ __align__(16) int8_t lds[];
    const CbArgs& p = ta.p;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int h = lane >> 5;
    const int c = lane & 31;
    const int tsl = wv >> 2;         // time slot
    const int wc = (wv & 3) % FS;    // channel of the fscrunch group
    const int wb = (wv & 3) / FS;    // beam group within the work-group
    constexpr int NWB = 4 / FS;

    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int bgroup = logical % ta.n_bgroups;
    const int rest = logical / ta.n_bgroups;
    const int tgroup = rest % ta.n_tgroups;
    const int fo = rest / ta.n_tgroups;
    if (fo >= p.Fo) return;   // uniform over the work-group

    const int K = 2 * p.A;
    const int f = fo * FS + wc;
    const int beam_base = (bgroup * NWB + wb) * 16 * RT;
    const size_t TP = (size_t)2 * p.nsamples;
    const int slot_tiles = ta.npasses * ta.ntp;
    const int tile0 = (tgroup * NTS + tsl) * slot_tiles;

    // Start-up loads first, all in flight together: the channel's quantiser
    // thresholds (to LDS below), the weights, then the first staged tiles.
    // One threshold per thread, two for the first threads of a one-slot
    // work-group (256 threads, kThr = 257).
    static_assert(512 * NTS >= kThr, "at most two thresholds per thread");
    const long long thr_v = tid < kThr ? p.thr[(size_t)fo * kThr + tid] : 0;
    const long long thr_v2 = tid + 256 * NTS < kThr ? p.thr[(size_t)fo * kThr + tid + 256 * NTS] : 0;

    // Weights (B operand).  Tile r holds the Re (r even) or Im (r odd)
    // embedding rows of the 32 beams beam_base + 32 (r / 2) + n, n = lane %
    // 32, so a lane's Re and Im outputs belong to one beam and its power
    // sums in-lane.
    v4i w[RT][KSTEPS];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int b = beam_base + 32 * g + c;
        const bool bvalid = b < p.B;
        const int8_t* wcol = p.fba + ((size_t)f * p.B + (bvalid ? b : p.B - 1)) * K;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) w[2 * g][s] = load_wslice<true>(wcol, 16 * (h * KSTEPS + s), K, bvalid);
    }

    // Staging.  A wave's 16-byte load lands at LDS offset 16 * lane of a
    // 1 KiB run.  FTPA: the bank swizzle is chosen on the global side:
    // physical piece pp of column col holds K piece (pp + rot(col)) % PR with
    // rot(col) = col * KPAD / 256.  K pieces past the column (k >= K) re-read
    // piece 0: the weights there are zero.  TAFTP: a tile is 16 * KSTEPS
    // antenna rows of 64 bytes (its 32 columns, contiguous in TAFTP), and a
    // load fills 16 rows, 4 lanes a row; rows past A re-read antenna 0.
    const int kpieces = K / 16;
    // LDS: the channel's thresholds first (at address 0, so that a
    // threshold read takes the table's offset as its immediate), then the
    // ring, the dummy run and the fscrunch reduction.
    long long* thr = reinterpret_cast<long long*>(lds);                   // [kThr] of channel fo
    int8_t* const ring = lds + kThrBytes;
    int8_t* dummy = ring + NB * St::SET;
    long long* red = reinterpret_cast<long long*>(dummy + St::DUMMY);   // [waves][VALS][64]
    // A set holds one tile per (time slot, channel): tile index
    // tsl * FS + ch.  Load m of this wave fills 1 KiB of tile ti = u / KSTEPS
    // (u = wave + NW * m, the same tile for the whole wave): the channel's
    // base is wave-uniform (scalar), and each lane adds its column and
    // swizzled K piece as a 32-bit offset.
    const int8_t* cbase[St::M];
    uint32_t goff[St::M];
    int gtile[St::M];
    int ldst[St::M];
#pragma unroll
    for (int m = 0; m < St::M; ++m) {
        const int u = wv + St::NW * m;
        if (u < St::NI) {
            const int ti = u / KSTEPS;
            if constexpr (TA) {
                const int kp = u - ti * KSTEPS;
                int an = 16 * kp + (lane >> 2);
                if (an >= p.A) an = 0;
                cbase[m] = p.ftpa + (size_t)(fo * FS + ti % FS) * p.S * 4;
                goff[m] = (uint32_t)(p.a0 + an) * (uint32_t)(p.F * p.S * 4) + 16 * (lane & 3);
            } else {
                const int rem = (u - ti * KSTEPS) * 1024 + lane * 16;
                const int col = rem / KPAD;
                const int pp = (rem - col * KPAD) >> 4;
                const int j = (pp + ((col * KPAD) >> 8)) % PR;
                const int jj = j < kpieces ? j : 0;
                cbase[m] = p.ftpa + (size_t)(fo * FS + ti % FS) * TP * p.vstride + p.voff;
                goff[m] = (uint32_t)(col * p.vstride + 16 * jj);
            }
            gtile[m] = (tgroup * NTS + ti / FS) * slot_tiles;
            ldst[m] = u * 1024;
        } else {
            cbase[m] = TA ? p.ftpa : p.ftpa + (size_t)(fo * FS) * TP * p.vstride + p.voff;
            goff[m] = 0;
            gtile[m] = 0;
            ldst[m] = -1;
        }
    }
    const size_t tile_stride = (size_t)32 * p.vstride;
    const int tmask = (1 << ta.lgtph) - 1;
    // TAFTP: each load keeps a running 32-bit offset of its next tile (the
    // block is below 4 GiB, fbf_cb_beamform_taftp's condition), since the
    // issues take consecutive tiles: 64 bytes on within a heap, the next
    // heap's first tile at a heap boundary; loads past the last tile keep
    // re-reading it.  A handful of scalar instructions per load instead of a
    // 64-bit tile address (every wave of a compute unit shares one scalar
    // unit, and the loads are issued right after a barrier).
    uint32_t toff_run[St::M];
    int tile_run[St::M];
    if constexpr (TA) {
#pragma unroll
        for (int m = 0; m < St::M; ++m) {
            const int t = gtile[m] < ta.ntiles ? gtile[m] : ta.ntiles - 1;
            tile_run[m] = gtile[m];
            toff_run[m] = (uint32_t)(t >> ta.lgtph) * (uint32_t)ta.hstride + ((uint32_t)(t & tmask) << 6);
            cbase[m] = uniform_ptr(cbase[m]);
        }
    }
    const uint32_t heap_jump = (uint32_t)ta.hstride - (uint32_t)tmask * 64u;
    auto issue = [&](int q, int slot) {
        int8_t* set = ring + slot * St::SET;
#pragma unroll
        for (int m = 0; m < St::M; ++m) {
            int8_t* dst = ldst[m] >= 0 ? set + ldst[m] : dummy + wv * 1024;
            if constexpr (TA) {
                dma16(cbase[m] + toff_run[m] + goff[m], dst);
                ++tile_run[m];
                if (tile_run[m] < ta.ntiles) toff_run[m] += (tile_run[m] & tmask) ? 64u : heap_jump;
            } else {
                const int tt = gtile[m] + q < ta.ntiles ? gtile[m] + q : ta.ntiles - 1;
                dma16(uniform_ptr(cbase[m] + (size_t)tt * tile_stride) + goff[m], dst);
            }
        }
    };
    // Physical offsets of this lane's K pieces in its column (FTPA), or of
    // the row piece this lane reads for K slice s (TAFTP, tr_operands): rows
    // 8 (h KSTEPS + s) + (lane / 4) % 4, columns 16 ((lane / 16) % 2) +
    // 4 (lane % 4), so lane c + 32 h gets the antennas of K piece
    // h KSTEPS + s, as FTPA's column holds them.
    int aoff[KSTEPS];
    if constexpr (TA) {
        const int rr = (lane >> 2) & 3, xx = lane & 3, gh = (lane >> 4) & 1;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s)
            aoff[s] = (tsl * FS + wc) * KSTEPS * 1024 + (8 * (h * KSTEPS + s) + rr) * 64 + 32 * gh + 8 * xx;
    } else {
        const int rot = (c * KPAD) >> 8;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s)
            aoff[s] = ((tsl * FS + wc) * 32 + c) * KPAD + 16 * ((h * KSTEPS + s - rot + 32 * PR) % PR);
    }

    const bool wave_active = beam_base < p.B;
#pragma unroll
    for (int q = 0; q < NB - St::G; ++q) issue(q, q);
    int step = 0;
    int rs = 0;              // ring slot read this step
    int is = NB - St::G;     // next ring slot to fill

    bool m128 = false;
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            m128 |= raw_has_m128(w[2 * g][s]);
            w[2 * g + 1][s] = embed_slice(w[2 * g][s], true);
            w[2 * g][s] = embed_slice(w[2 * g][s], false);
        }
    // A work-group any of whose waves holds a weight wi = -128 computes its
    // outputs on the exact path instead.  The flags pass through the
    // thresholds' LDS before the thresholds are written there: written, a
    // barrier, read by every wave, a barrier.
    {
        int* const flags = reinterpret_cast<int*>(thr);
        const bool wave_m128 = wave_any(m128);   // every lane votes (not inside the lane-0 branch)
        if (lane == 0) flags[wv] = wave_m128 ? 1 : 0;
        wait_vm_lgkm0<63>();
        block_barrier();
        int any = 0;
#pragma unroll
        for (int i = 0; i < St::NW; ++i) any |= flags[i];
        any = __builtin_amdgcn_readfirstlane(any);
        wait_vm_lgkm0<63>();
        block_barrier();
        if (any) {
            wait_vm_lgkm0<0>();   // the staging loads land before LDS is released
            const int tend = (tgroup + 1) * NTS * slot_tiles;
            cb_exact_outputs<MODE_INT8>(p, fo, bgroup * NWB * 16 * RT, NWB * 16 * RT,
                                        (long long)tgroup * NTS * slot_tiles * SPT / ta.tps,
                                        (long long)(tend < ta.ntiles ? tend : ta.ntiles) * SPT / ta.tps, tid,
                                        (int)blockDim.x);
            if constexpr (IBF) {
                if (bgroup == 0)
                    cb_exact_ib_partials(p, ta.ibp, ta.ntiles, fo, tgroup * NTS * slot_tiles,
                                         tend < ta.ntiles ? tend : ta.ntiles, tid, (int)blockDim.x);
            }
            return;
        }
    }

    // Quantiser: the thresholds are visible after the tile loop's first
    // barrier.
    const float q_off = p.offsets[fo], q_sc = p.scales[fo];
    const bool q_thr = thr_applies(q_off, q_sc);
    const bool q_f32 = q_thr && thr_f32_applies(q_off, q_sc);
    const QuantConst qz = quant_const(q_off, q_sc);
    if (tid < kThr) thr[tid] = thr_v;
    if (tid + 256 * NTS < kThr) thr[tid + 256 * NTS] = thr_v2;
    const long long* const thr_mid = thr + 128;
    // Quantise and store one pass's totals (staged [fo][to][b] int8); with
    // FS > 1 the waves of a beam group share the totals out.
    // Value v of a lane: slot 2v + h of beam c (RT = 2), else slot
    // v % NSLOT of beam 64 (v / NSLOT) + 32 h + c.  The row of slot 2v
    // (RT = 2) or v % NSLOT is wave-uniform, so its address is scalar and
    // each lane adds a 32-bit offset (the next row for the upper half when
    // RT = 2).
    int8_t* const out_f = static_cast<int8_t*>(p.out) + (size_t)fo * p.nsamples_out * p.B;
    const int b_lane = RT == 2 ? beam_base + c : beam_base + 32 * h + c;
    const int lane_off = (RT == 2 ? h * p.B : 0) + b_lane;
    // Output windows number below 2^31 (cb_tr_applies), so window indices
    // compare as int: a scalar compare for the wave-uniform rows.
    const int nout = (int)p.nsamples_out;
    const bool all_beams = wave_active && beam_base + 16 * RT <= p.B;
    // All of a pass's values are estimated first, then their thresholds
    // read (the LDS reads in flight together, one wait), then stored.
    auto store_pass = [&](const long long (&t)[VALS], int w0) {
        int q[VALS];   // (ints: no packing into dwords before the byte stores)
        if (q_thr) {
            int e[VALS];
            if (q_f32) {
#pragma unroll
                for (int v = 0; v < VALS; ++v) e[v] = thr_estimate_f32(t[v], qz);
            } else {
#pragma unroll
                for (int v = 0; v < VALS; ++v) e[v] = thr_estimate(t[v], qz);
            }
#pragma unroll
            for (int v = 0; v < VALS; ++v) q[v] = thr_settle(t[v], e[v], thr_mid);
        } else {
#pragma unroll
            for (int v = 0; v < VALS; ++v) q[v] = quantise((double)t[v], q_off, q_sc, p.quant);
        }
        // Every beam of the wave and every window of the pass in range (the
        // usual case, uniform over the wave): the four rows' bases are
        // scalar and the stores need no per-store checks.
        if constexpr (RT != 2 && FS == 1) {
            if (all_beams && w0 + NSLOT <= nout) {
                int8_t* const base = out_f + (size_t)w0 * p.B;
#pragma unroll
                for (int v = 0; v < VALS; ++v)
                    global_ptr(uniform_ptr(base + (size_t)(v % NSLOT) * p.B + 64 * (v / NSLOT)))[(unsigned)lane_off] =
                        q[v];
                return;
            }
        }
#pragma unroll
        for (int v = 0; v < VALS; ++v) {
            if (FS > 1 && v % FS != wc) continue;
            const int row = w0 + (RT == 2 ? 2 * v : v % NSLOT);
            const int bsh = RT == 2 ? 0 : 64 * (v / NSLOT);
            // Uniform bounds first (a scalar branch), then the lane's.
            if (!wave_active || row >= nout) continue;
            if (b_lane + bsh < p.B && (RT != 2 || row + h < nout))
                global_ptr(uniform_ptr(out_f + (size_t)row * p.B + bsh))[(unsigned)lane_off] = q[v];
        }
    };
    const int prio_phase = blockIdx.x >= gridDim.x / 2 ? 1 : 0;
    // First window of the pass; tile0 is a multiple of the slot's tiles, so
    // the division is exact.
    int w0_pass = tile0 * SPT / ta.tps;
    for (int pass = 0; pass < ta.npasses; ++pass) {
        long long pw[NG][NSLOT];
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) pw[g][sl] = 0;
        // Passes past the end of the channel (the last slots of a one-round
        // grid) still take part in the staging and the barriers but leave
        // the MFMA pipe to the other waves of the SIMD.
        const bool tiles_live = tile0 + pass * ta.ntp < ta.ntiles;
        const bool live = wave_active && tiles_live;
        // Incoherent-beam partials: in beam group 0, one wave of each time
        // slot per step (whether it holds beams or not) forms the tile's.
        // The role rotates with the step (and, with two time slots, the
        // slot), so that a wave issues one partials store per four steps
        // (stores count against the staging loads' vmcnt wait) and the
        // SIMDs take turns.
        const bool ib_live = IBF && bgroup == 0 && tiles_live;

        // Tiles of the pass: NSLOT/SPT groups of tps tiles (tps = 1 unless
        // a window spans several tiles).
#pragma unroll
        for (int g = 0; g < NSLOT / SPT; ++g) {
            for (int x = 0; x < ta.tps; ++x) {
                // Every G steps: this wave's loads of the next G sets are
                // done once WAIT later loads remain; the barrier makes every
                // wave's loads visible and frees the G slots read since the
                // last barrier, which are refilled NB - G steps ahead.
                // One-slot work-groups (two per compute unit, one wave each
                // per SIMD): the issue arbiter favours the older wave of a
                // SIMD, so the two work-groups take turns at a raised
                // priority, every 16 steps in opposite phase (the halves of
                // the grid), instead of the older one finishing first and
                // leaving its SIMDs one wave for the rest.
                if constexpr (NTS == 1) {
                    if (step % 16 == 0) {
                        if (((step >> 4) + prio_phase) & 1) __builtin_amdgcn_s_setprio(2);
                        else __builtin_amdgcn_s_setprio(0);
                    }
                }
                if (step % St::G == 0) {
                    wait_vm_lgkm0<St::WAIT>();
                    block_barrier();
#pragma unroll
                    for (int j = 0; j < St::G; ++j) {
                        issue(step + NB - St::G + j, is);
                        is = is + 1 == NB ? 0 : is + 1;
                    }
                }
                const bool ib_now = ib_live && ((step + 2 * tsl) & 3) == (wv & 3);
                if (live || ib_now) {
                    const int8_t* src = ring + rs * St::SET;
                    v4i a[KSTEPS];
                    if constexpr (TA) {
                        tr_operands<KSTEPS>(src + aoff[0], a);
                    } else {
#pragma unroll
                        for (int s = 0; s < KSTEPS; ++s) a[s] = *reinterpret_cast<const v4i*>(src + aoff[s]);
                    }
                    if (live) detect_step<RT, KSTEPS, SPT, NSLOT>(a, w, pw, g);
                    if constexpr (IBF) {
                        if (ib_now) ib_tile<KSTEPS>(a, ta.ibp, ta.ntiles, p.A, fo, tile0 + pass * ta.ntp + g * ta.tps + x,
                                                    lane);
                    }
                }
                rs = rs + 1 == NB ? 0 : rs + 1;
                ++step;
            }
        }

        // A window's power is split over the row halves (lane and lane ^ 32):
        // one v_permlane32_swap reduce-scatter per value, lanes below 32
        // keeping the even beam group of each pair (RT >= 4) or the even
        // slots (RT = 2).
        long long tot[VALS];
        if constexpr (RT == 2) {
#pragma unroll
            for (int v = 0; v < VALS; ++v) tot[v] = swap32_sum(pw[0][2 * v], pw[0][2 * v + 1]);
        } else {
#pragma unroll
            for (int gp = 0; gp < NG / 2; ++gp)
#pragma unroll
                for (int sl = 0; sl < NSLOT; ++sl)
                    tot[gp * NSLOT + sl] = swap32_sum(pw[2 * gp][sl], pw[2 * gp + 1][sl]);
        }

        // Sum over the fscrunch channels (one per wave of a beam group).
        if constexpr (FS > 1) {
#pragma unroll
            for (int v = 0; v < VALS; ++v) red[(wv * VALS + v) * 64 + lane] = tot[v];
            wait_vm_lgkm0<63>();
            block_barrier();
#pragma unroll
            for (int v = 0; v < VALS; ++v) {
                long long sum = 0;
#pragma unroll
                for (int ch = 0; ch < FS; ++ch) sum += red[((tsl * 4 + wb * FS + ch) * VALS + v) * 64 + lane];
                tot[v] = sum;
            }
            wait_vm_lgkm0<63>();
            block_barrier();
        }

        store_pass(tot, w0_pass);
        w0_pass += NSLOT;   // a pass holds NSLOT windows (ta.ntp = NSLOT / SPT * tps tiles)
    }
    // The loads past the last tile are still in flight: LDS must not be
    // released before they land.
    wait_vm_lgkm0<0>();
}

// Launch of cb_tr_k.  Supports 8 | A (16-byte K slices), A <= 128,
// tscrunch a power of two >= 4 (window 8, 16, 32 or a multiple of 32
// columns), fscrunch 1, 2 or 4 with fscrunch * ceil(A / 16) <= 8, int8
// output.
// Windows per pass.  With 64 beams per wave and windows of 16 or more
// columns, two windows per pass keep a wave within 128 VGPRs, so four waves
// share a SIMD (tr_nts).  Windows of 8 columns fill a tile four times, so a
// pass holds at least 4.
inline int cb_tr_nslot(int rt, int W)
{
    if (W == 16 && rt == 4) return 2;
    if (W <= 16) return rt == 2 ? 8 : 4;
    return W == 32 && rt != 4 ? 4 : 2;
}

inline int cb_tr_nb(int ni) { return 60 / ni < 2 ? 2 : (60 / ni > 8 ? 8 : 60 / ni); }

// Time slots per work-group of the transposed kernel: four with 64 beams
// per wave and two windows per pass (at most 128 VGPRs a wave, four waves per
// SIMD), one with 128 beams per wave (256 VGPRs, two work-groups and so two
// waves per SIMD), else three.  With four waves a SIMD overlaps one wave's
// MFMAs with the others' detection better: at 64 antennas the VALU work per
// step outweighs the MFMA work.  With 128 beams per wave, one work-group of
// two time slots kept its two waves of a SIMD in lockstep, and the older one
// waited at every barrier for the younger; two one-slot work-groups with
// alternating priority measured 2.0-2.4 % faster.
constexpr int tr_nts(int rt, int nslot) { return rt == 8 ? 1 : (rt == 4 && nslot == 2 ? 4 : 3); }

inline size_t cb_tr_shm(int fs, int ksteps, int vals, int nts)
{
    const int nw = 4 * nts;
    const int ni = nts * fs * ksteps;
    return (size_t)cb_tr_nb(ni) * ni * 1024 + ((ni % nw) ? (size_t)nw * 1024 : 0) +
           (fs > 1 ? (size_t)nw * vals * 64 * sizeof(long long) : 0) + kThrBytes;
}

inline bool cb_tr_supported(int A, int ts, int fs, int rt)
{
    if (A % 8 != 0 || A > 128) return false;
    if (ts < 4 || (ts & (ts - 1)) != 0) return false;
    if (!(fs == 1 || fs == 2 || fs == 4)) return false;
    const int ksteps = (2 * A + 31) / 32;
    if (rt == 8 ? (fs != 1 || ksteps > 4) : rt * ksteps > 16) return false;
    if (fs * ksteps > 8) return false;
    const int nslot = cb_tr_nslot(rt, 2 * ts);
    const int vals = rt == 2 ? nslot / 2 : (rt / 4) * nslot;
    return cb_tr_shm(fs, ksteps, vals, tr_nts(rt, nslot)) <= 65536;
}

// Compute units of the current device (cached per device).
inline int device_cu_count()
{
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

// One round of resident work-groups: every (channel group, beam group) gets
// (resident work-groups) / (channel groups x beam groups) work-groups, each a
// contiguous run of passes, so the per-work-group start-up (weights, the
// first staged tiles) is paid once per resident slot and all work-groups end
// together.
template <int RT, int KSTEPS, int SPT, int NSLOT, int FS, bool TA, bool IBF = false>
hipError_t launch_tr_kt(CbTrArgs a, int passes, hipStream_t s)
{
    constexpr int NTS = tr_nts(RT, NSLOT);
    const size_t shm = cb_tr_shm(FS, KSTEPS, TrRed<RT, NSLOT>::VALS, NTS);
    static int occ = 0;
    if (occ == 0) {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, cb_tr_k<RT, KSTEPS, SPT, NSLOT, FS, NTS, TA, IBF>, 256 * NTS,
                                                         shm) != hipSuccess ||
            o < 1)
            o = 1;
        occ = o;
    }
    // Work-groups per (channel group, beam group): one round of resident
    // work-groups where the problem allows; each covers NTS slots of
    // npasses passes.
    const long long units = (long long)a.p.Fo * a.n_bgroups;
    long long per_unit = (long long)occ * device_cu_count() / units;
    if (per_unit < 1) per_unit = 1;
    const long long max_groups = (passes + NTS - 1) / NTS;
    if (per_unit > max_groups) per_unit = max_groups;
    a.npasses = (int)((passes + per_unit * NTS - 1) / (per_unit * NTS));
    a.n_tgroups = (int)((passes + (long long)a.npasses * NTS - 1) / ((long long)a.npasses * NTS));
    const long long nblocks = units * a.n_tgroups;
    if (nblocks <= 0 || nblocks >= (1ll << 31)) return hipErrorInvalidConfiguration;
    cb_tr_k<RT, KSTEPS, SPT, NSLOT, FS, NTS, TA, IBF><<<dim3((unsigned)nblocks), 256 * NTS, shm, s>>>(a);
    return hipGetLastError();
}

// TAFTP staging (a.p.S > 0) is built for 64 and 128 beams per wave and
// fscrunch 1 only (cb_taftp_direct): with 64 beams per wave no fscrunch above
// 1 fits the kernel's LDS (cb_tr_supported).  The incoherent-beam partials
// (a.ibp) are built for 128 beams per wave (cb_ib_fused).
template <int RT, int KSTEPS, int SPT, int NSLOT, int FS>
hipError_t launch_tr_k(CbTrArgs a, int passes, hipStream_t s)
{
    if (a.p.S > 0) {
        if constexpr (RT != 2 && FS == 1) {
            if (a.ibp) {
                if constexpr (RT == 8) return launch_tr_kt<RT, KSTEPS, SPT, NSLOT, FS, true, true>(a, passes, s);
                return hipErrorInvalidValue;
            }
            return launch_tr_kt<RT, KSTEPS, SPT, NSLOT, FS, true>(a, passes, s);
        }
        return hipErrorInvalidValue;
    }
    return launch_tr_kt<RT, KSTEPS, SPT, NSLOT, FS, false>(a, passes, s);
}

template <int RT, int SPT, int NSLOT, int FS>
hipError_t launch_tr_ks(const CbTrArgs& a, int ksteps, int passes, hipStream_t s)
{
    if constexpr (FS * 2 <= 8) {
        if (ksteps == 1) return launch_tr_k<RT, 1, SPT, NSLOT, FS>(a, passes, s);
        if (ksteps == 2) return launch_tr_k<RT, 2, SPT, NSLOT, FS>(a, passes, s);
    }
    if constexpr (FS * 4 <= 8) {
        if (ksteps == 3) return launch_tr_k<RT, 3, SPT, NSLOT, FS>(a, passes, s);
        if (ksteps == 4) return launch_tr_k<RT, 4, SPT, NSLOT, FS>(a, passes, s);
    }
    if constexpr (FS == 1 && RT == 2) {
        switch (ksteps) {
        case 5: return launch_tr_k<RT, 5, SPT, NSLOT, FS>(a, passes, s);
        case 6: return launch_tr_k<RT, 6, SPT, NSLOT, FS>(a, passes, s);
        case 7: return launch_tr_k<RT, 7, SPT, NSLOT, FS>(a, passes, s);
        case 8: return launch_tr_k<RT, 8, SPT, NSLOT, FS>(a, passes, s);
        default: break;
        }
    }
    return hipErrorInvalidValue;
}

template <int RT, int FS>
hipError_t launch_tr_w(const CbTrArgs& a, int W, int ksteps, int passes, hipStream_t s)
{
    // Windows per pass as cb_tr_nslot
    constexpr int small = RT == 2 ? 8 : 4;
    if (W == 8) return launch_tr_ks<RT, 4, small, FS>(a, ksteps, passes, s);
    if (W == 16) return launch_tr_ks<RT, 2, RT == 4 ? 2 : small, FS>(a, ksteps, passes, s);
    if (W == 32) return launch_tr_ks<RT, 1, RT == 4 ? 2 : 4, FS>(a, ksteps, passes, s);
    return launch_tr_ks<RT, 1, 2, FS>(a, ksteps, passes, s);
}

template <int RT>
hipError_t launch_tr(const CbTrArgs& a, int fs, int W, int ksteps, int passes, hipStream_t s)
{
    if constexpr (RT == 8) {
        if (fs != 1 || ksteps > 4) return hipErrorInvalidValue;
        return launch_tr_w<RT, 1>(a, W, ksteps, passes, s);
    } else {
        if (fs == 1) return launch_tr_w<RT, 1>(a, W, ksteps, passes, s);
        if (fs == 2) return launch_tr_w<RT, 2>(a, W, ksteps, passes, s);
        return launch_tr_w<RT, 4>(a, W, ksteps, passes, s);
    }
}

// ---------------------------------------------------------------------------
// Wave-private staging (cb_ws_k)
//
// The contraction, detection, reduce-scatter and quantiser of cb_tr_k with
// RT = 4 (a wave holds 64 beams, Re and Im tiles) and fscrunch 1, but every
// wave stages its own voltage tiles in a private LDS ring and waits for
// nothing but its own loads: no work-group barrier in the tile loop, so the
// waves of a SIMD drift apart and one wave's MFMAs overlap another wave's
// detection (the 64-bit multiply-adds take as long as the MFMAs at 64
// antennas).  Work is dealt dynamically.  A unit is (output channel, 64-beam
// group, time range); a work-group of 12 waves serves U units, the waves
// starting 12 / U to a unit; a wave takes chunks (whole passes, at least a
// ring's worth of tiles) from its unit's LDS counter and, when the unit runs
// dry, moves to the next unit of the work-group (reloading its weights), so
// the waves of a work-group end within a chunk of each other.
//
// Ring: NB tiles per wave; the tile at sequence position q sits in slot
// q % NB.  Step q waits for its own loads of tile q (a fixed s_waitcnt:
// (NB - 1) tiles issued after it), reads the tile into registers, and then
// reuses the slot for tile q + NB.  A slot is written only after the same
// wave's reads of it have completed, and the loads are issued for every
// step (past the end of the work they re-read a valid tile), so the count
// stays exact.
// ---------------------------------------------------------------------------
// Returns the old value of the LDS word at byte address addr and adds v.
// Written as one asm block that also waits for the result: the compiler
// puts a full vmcnt(0) wait before its own LDS atomics while global->LDS
// loads are in flight, which would drain a wave's staging ring at every
// chunk.
__device__ __forceinline__ int lds_fetch_add(unsigned addr, int v)
{
    int old;
    __asm__ volatile("ds_add_rtn_u32 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)" : "=v"(old) : "v"(addr), "v"(v) : "memory");
    return old;
}

constexpr int kWsWaves = 12;
constexpr int kWsRingBudget = 144 * 1024;
constexpr int kWsMaxUnits = 6;

template <int KSTEPS>
struct WsStage {
    static constexpr int KPAD = 32 * KSTEPS;
    static constexpr int TILE = 32 * KPAD;                  // bytes per tile
    static constexpr int LPT = KSTEPS;                      // 1 KiB loads per tile
    static constexpr int NB0 = kWsRingBudget / (kWsWaves * TILE);
    static constexpr int NB = NB0 < 2 ? 2 : (NB0 > 8 ? 8 : NB0);
    static constexpr int WAIT = (NB - 1) * LPT;
};

inline int cb_ws_nb(int ksteps)
{
    const int nb = kWsRingBudget / (kWsWaves * 1024 * ksteps);
    return nb < 2 ? 2 : (nb > 8 ? 8 : nb);
}

struct CbWsArgs {
    CbArgs p;
    int ntiles;      // 32-column tiles per channel
    int tps;         // tiles per window (W >= 32), else 1
    int ntp;         // tiles per pass
    int ct;          // tiles per chunk: a whole number of passes, >= NB
    int nbg;         // 64-beam groups
    int ntr;         // time ranges per (channel, beam group)
    int tr_chunks;   // chunks per time range
    int nunits;      // Fo * nbg * ntr
    int U;           // units per work-group
};

template <int KSTEPS, int SPT, int NSLOT>
__global__ __launch_bounds__(64 * kWsWaves) void cb_ws_k(CbWsArgs wa)
{
    using St = WsStage<KSTEPS>;
    constexpr int KPAD = St::KPAD;
    constexpr int PR = 2 * KSTEPS;
    constexpr int NB = St::NB;
    constexpr int RT = 4, NG = 2;
    static_assert(NSLOT % SPT == 0, "slots");
    static_assert(St::WAIT <= 63, "vmcnt");
    extern __shared__ __align__(16) int8_t lds[];
    const CbArgs& p = wa.p;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = lane >> 5;
    const int c = lane & 31;
    const int U = wa.U;
    const int unit0 = xcd_remap(blockIdx.x, gridDim.x) * U;
    int8_t* const ring = lds + wv * NB * St::TILE;
    long long* const thr_all = reinterpret_cast<long long*>(lds + kWsWaves * NB * St::TILE);   // [U][kThr]
    int* const counter = reinterpret_cast<int*>(thr_all + U * kThr);                          // [U]
    int* const m128_flags = counter + U;                                                      // [kWsWaves]

    // Units: id = ((fo * ntr + tr) * nbg + bg).
    auto unit_fo = [&](int gu) { return gu / (wa.ntr * wa.nbg); };
    for (int i = tid; i < U * kThr; i += 64 * kWsWaves) {
        const int gu = unit0 + i / kThr;
        thr_all[i] = gu < wa.nunits ? p.thr[(size_t)unit_fo(gu) * kThr + i % kThr] : 0;
    }
    if (tid < U) counter[tid] = 0;
    // Weights of every unit of the work-group: one holding a wi = -128 sends
    // the work-group to the exact path (before any staging load is issued).
    const int K = 2 * p.A;
    {
        const int kp = K / 16;
        const int per_unit = 64 * kp;
        bool m128 = false;
        for (int i = tid; i < U * per_unit; i += 64 * kWsWaves) {
            const int gu = unit0 + i / per_unit;
            const int r = i % per_unit;
            const int b = (gu % wa.nbg) * 64 + r / kp;
            if (gu < wa.nunits && b < p.B)
                m128 |= raw_has_m128(
                    *reinterpret_cast<const v4i*>(p.fba + ((size_t)unit_fo(gu) * p.B + b) * K + 16 * (r % kp)));
        }
        const bool wave_m128 = wave_any(m128);   // every lane votes (not inside the lane-0 branch)
        if (lane == 0) m128_flags[wv] = wave_m128 ? 1 : 0;
    }
    __syncthreads();
    {
        int any = 0;
#pragma unroll
        for (int i = 0; i < kWsWaves; ++i) any |= m128_flags[i];
        if (__builtin_amdgcn_readfirstlane(any)) {
            for (int u = 0; u < U; ++u) {
                const int gu = unit0 + u;
                if (gu >= wa.nunits) break;
                const int tr = (gu / wa.nbg) % wa.ntr;
                const long long t0 = (long long)tr * wa.tr_chunks * wa.ct;
                long long t1 = t0 + (long long)wa.tr_chunks * wa.ct;
                if (t1 > wa.ntiles) t1 = wa.ntiles;
                cb_exact_outputs<MODE_INT8>(p, unit_fo(gu), (gu % wa.nbg) * 64, 64, t0 * SPT / wa.tps,
                                            t1 * SPT / wa.tps, tid, (int)blockDim.x);
            }
            return;
        }
    }

    const size_t TP = (size_t)2 * p.nsamples;
    const int kpieces = K / 16;
    const int chunk_tiles = wa.ct;

    // This lane's part of a tile load m: column and K piece (swizzled as in
    // cb_tr_k: physical piece pp of column col holds K piece
    // (pp + col * KPAD / 256) % PR; pieces past K re-read piece 0, whose
    // weights there are zero).
    int goff[St::LPT];
#pragma unroll
    for (int m = 0; m < St::LPT; ++m) {
        const int off = m * 1024 + lane * 16;
        const int col = off / KPAD;
        const int pp = (off - col * KPAD) >> 4;
        const int j = (pp + ((col * KPAD) >> 8)) % PR;
        goff[m] = col * p.vstride + 16 * (j < kpieces ? j : 0);
    }
    int aoff[KSTEPS];
    {
        const int rot = (c * KPAD) >> 8;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) aoff[s] = c * KPAD + 16 * ((h * KSTEPS + s - rot + 32 * PR) % PR);
    }

    // Chunk descriptors: (local unit << 24) | chunk, or -1 once every unit
    // of the work-group is dry.
    int home = (wv * U) / kWsWaves;
    const unsigned counter_lds = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)counter;
    auto grab = [&]() -> int {
        for (int k = 0; k < U; ++k) {
            int got = 0;
            if (lane == 0) got = lds_fetch_add(counter_lds + 4u * (unsigned)home, 1);
            got = __builtin_amdgcn_readfirstlane(got);
            const int gu = unit0 + home;
            const int tr = (gu / wa.nbg) % wa.ntr;
            if (gu < wa.nunits && got < wa.tr_chunks && (tr * wa.tr_chunks + got) * chunk_tiles < wa.ntiles)
                return (home << 24) | got;
            home = home + 1 == U ? 0 : home + 1;
        }
        return -1;
    };
    // First tile (in its channel) and channel byte offset of a chunk.
    auto chunk_tile0 = [&](int d) {
        const int gu = unit0 + (d >> 24);
        const int tr = (gu / wa.nbg) % wa.ntr;
        return (tr * wa.tr_chunks + (d & 0xFFFFFF)) * chunk_tiles;
    };
    // The issue cursor's chunk as a first-tile address and a count of tiles
    // inside the channel (later tiles re-read the last one; past the end of
    // the work every load re-reads tile 0 of channel 0).
    const size_t tile_bytes = (size_t)32 * p.vstride;
    const int8_t* isrc = p.ftpa + p.voff;
    int ilast = 0;
    auto set_issue_chunk = [&](int d) {
        isrc = p.ftpa + p.voff;
        ilast = 0;
        if (d >= 0) {
            const int t0 = chunk_tile0(d);
            isrc += (size_t)unit_fo(unit0 + (d >> 24)) * TP * p.vstride + (size_t)t0 * tile_bytes;
            ilast = (wa.ntiles - t0 < chunk_tiles ? wa.ntiles - t0 : chunk_tiles) - 1;
        }
    };
    auto issue = [&](int off, int slot) {
        const int8_t* src = isrc + (size_t)(off < ilast ? off : ilast) * tile_bytes;
#pragma unroll
        for (int m = 0; m < St::LPT; ++m) dma16(src + (uint32_t)goff[m], ring + slot * St::TILE + m * 1024);
    };

    int cdesc = grab();          // chunk of the compute cursor
    int ioff = 0;                // issue cursor's tile within its chunk
    int pend = -1;               // next chunk, taken by the issue cursor
    set_issue_chunk(cdesc);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        issue(ioff, q);
        ++ioff;
    }

    // Weights of the unit being computed (loaded when it changes).
    v4i w[RT][KSTEPS];
    int wunit = -1;
    int fo = 0, beam_base = 0;
    float q_off = 0.f, q_sc = 1.f;
    bool q_thr = false;
    QuantConst qz{0.0, 1.0, 1.0f, 0.0f};
    const long long* thr = thr_all;

    int q = 0;   // sequence position of the compute cursor
    while (cdesc >= 0) {
        const int u = cdesc >> 24;
        if (u != wunit) {
            wunit = u;
            const int gu = unit0 + u;
            fo = unit_fo(gu);
            beam_base = (gu % wa.nbg) * 64;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int b = beam_base + 32 * g + c;
                const bool bvalid = b < p.B;
                const int8_t* wcol = p.fba + ((size_t)fo * p.B + (bvalid ? b : p.B - 1)) * K;
#pragma unroll
                for (int s = 0; s < KSTEPS; ++s)
                    w[2 * g][s] = load_wslice<true>(wcol, 16 * (h * KSTEPS + s), K, bvalid);
            }
#pragma unroll
            for (int g = 0; g < NG; ++g)
#pragma unroll
                for (int s = 0; s < KSTEPS; ++s) {
                    w[2 * g + 1][s] = embed_slice(w[2 * g][s], true);
                    w[2 * g][s] = embed_slice(w[2 * g][s], false);
                }
            q_off = p.offsets[fo];
            q_sc = p.scales[fo];
            q_thr = thr_applies(q_off, q_sc);
            qz = quant_const(q_off, q_sc);
            thr = thr_all + u * kThr;
        }
        const int ctile0 = chunk_tile0(cdesc);
        const int b_out = beam_base + 32 * h + c;
        const bool wave_active = beam_base < p.B;

        for (int pc = 0; pc < chunk_tiles; pc += wa.ntp) {
            long long pw[NG][NSLOT];
#pragma unroll
            for (int g = 0; g < NG; ++g)
#pragma unroll
                for (int sl = 0; sl < NSLOT; ++sl) pw[g][sl] = 0;
#pragma unroll
            for (int g = 0; g < NSLOT / SPT; ++g) {
                for (int x = 0; x < wa.tps; ++x) {
                    const int slot = q % NB;
                    wait_vm_lgkm0<St::WAIT>();
                    const int8_t* src = ring + slot * St::TILE;
                    v4i a[KSTEPS];
#pragma unroll
                    for (int s = 0; s < KSTEPS; ++s) a[s] = *reinterpret_cast<const v4i*>(src + aoff[s]);
                    wait_vm_lgkm0<63>();
                    // The slot is free again: refill it NB positions ahead.
                    if (ioff == chunk_tiles) {
                        pend = grab();
                        set_issue_chunk(pend);
                        ioff = 0;
                    }
                    issue(ioff, slot);
                    ++ioff;
                    detect_step<RT, KSTEPS, SPT, NSLOT>(a, w, pw, g);
                    ++q;
                }
            }
            // Reduce-scatter over the row halves (lanes below 32 keep beam
            // group 0), quantise, store.
            long long tot[NSLOT];
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) tot[sl] = swap32_sum(pw[0][sl], pw[1][sl]);
            const int win0 = (ctile0 + pc) * SPT / wa.tps;
#pragma unroll
            for (int v = 0; v < NSLOT; ++v) {
                const long long to = (long long)win0 + v;
                if (!wave_active || b_out >= p.B || to >= p.nsamples_out) continue;
                int8_t* row = static_cast<int8_t*>(p.out) + ((size_t)fo * p.nsamples_out + (size_t)to) * p.B;
                row[b_out] = q_thr ? quantise_thr(tot[v], qz, thr + 128)
                                   : quantise((double)tot[v], q_off, q_sc, p.quant);
            }
        }
        cdesc = pend;
        pend = -1;
    }
    // Loads issued past the last chunk are still in flight.
    wait_vm_lgkm0<0>();
}

// Units, chunks and units per work-group of cb_ws_k for one launch.
template <int KSTEPS, int SPT, int NSLOT>
hipError_t launch_ws_k(CbWsArgs a, hipStream_t s)
{
    const int nb = WsStage<KSTEPS>::NB;
    a.ct = a.ntp * ((nb + a.ntp - 1) / a.ntp);
    const int chunks = (a.ntiles + a.ct - 1) / a.ct;
    const long long base_units = (long long)a.p.Fo * a.nbg;
    const int ncu = device_cu_count();
    int ntr = 1;
    if (base_units < ncu) ntr = (int)((ncu + base_units - 1) / base_units);
    if (ntr > chunks) ntr = chunks;
    a.tr_chunks = (chunks + ntr - 1) / ntr;
    a.ntr = (chunks + a.tr_chunks - 1) / a.tr_chunks;
    const long long nunits = base_units * a.ntr;
    if (nunits >= (1ll << 24)) return hipErrorInvalidConfiguration;
    a.nunits = (int)nunits;
    static const int kU[] = {1, 2, 3, 4, 6};
    a.U = kWsMaxUnits;
    for (int u : kU)
        if ((nunits + u - 1) / u <= ncu) {
            a.U = u;
            break;
        }
    const long long nblocks = (nunits + a.U - 1) / a.U;
    const size_t shm = (size_t)kWsWaves * nb * WsStage<KSTEPS>::TILE + (size_t)a.U * (kThr * sizeof(long long) + 4) +
                       kWsWaves * sizeof(int);
    cb_ws_k<KSTEPS, SPT, NSLOT><<<dim3((unsigned)nblocks), 64 * kWsWaves, shm, s>>>(a);
    return hipGetLastError();
}

template <int SPT, int NSLOT>
hipError_t launch_ws_ks(const CbWsArgs& a, int ksteps, hipStream_t s)
{
    switch (ksteps) {
    case 1: return launch_ws_k<1, SPT, NSLOT>(a, s);
    case 2: return launch_ws_k<2, SPT, NSLOT>(a, s);
    case 3: return launch_ws_k<3, SPT, NSLOT>(a, s);
    case 4: return launch_ws_k<4, SPT, NSLOT>(a, s);
    default: return hipErrorInvalidValue;
    }
}

// Windows per pass of cb_ws_k.
inline int cb_ws_nslot(int W) { return W <= 32 ? 4 : 2; }

inline hipError_t launch_ws(const CbWsArgs& a, int W, int ksteps, hipStream_t s)
{
    if (W == 8) return launch_ws_ks<4, 4>(a, ksteps, s);
    if (W == 16) return launch_ws_ks<2, 4>(a, ksteps, s);
    if (W == 32) return launch_ws_ks<1, 4>(a, ksteps, s);
    return launch_ws_ks<1, 2>(a, ksteps, s);
}

// cb_ws_k serves 8 | A <= 64, fscrunch 1, tscrunch a power of two >= 4.
inline bool cb_ws_supported(int A, int ts, int fs)
{
    return A % 8 == 0 && A <= 64 && fs == 1 && ts >= 4 && (ts & (ts - 1)) == 0;
}

// ---------------------------------------------------------------------------
// Staged [Fo][To][B] int8 -> TBTF [To/ohs][B][ohs][Fo]
//
// A work-group moves a tile of up to 128 beams x tt samples (within one
// output heap) x up to 64 channels.  Loads are 8-byte runs of beams (a tile
// row is a whole 128-byte line, so no line is fetched by two work-groups);
// the tile is scattered into LDS as [t][b][f] (channel bytes contiguous) and
// written out as 4-byte words of channels, so one wave stores whole rows of
// tt * Fo contiguous bytes per beam.  LDS dword layout: b pitch BP (odd),
// t pitch TP = 128 * BP + 16, so that the word reads of a wave (16 words of
// one (t, b) row x 4 samples) fall in distinct banks.
// ---------------------------------------------------------------------------
constexpr int kTrBeams = 128;
constexpr int kTrMaxChans = 64;
constexpr int kTrLdsBudget = 40 * 1024;

struct TbtfArgs {
    const int8_t* src;
    int8_t* dst;
    long long To;
    int B, Fo, ohs;
    int tt, ft, bp, tp;           // tile samples / channels, LDS pitches (dwords)
    int n_ttiles, n_ftiles, n_btiles;
};

// The first wj.nblocks work-groups generate weights (the next block's,
// weights_job_run), the next job.nblocks finish the incoherent beam from the
// coherent kernel's partials (ib_finish_run), both beside the reorder; the
// rest reorder, numbered from 0.
__global__ __launch_bounds__(256) void tbtf_k(TbtfArgs a, IbFinishJob job, WeightsJob wj)
{
    extern __shared__ __align__(16) int8_t tr_lds[];
    if (weights_job_run(wj, threadIdx.x)) return;   // uniform over the work-group
    const int b1 = (int)blockIdx.x - wj.nblocks;
    if (b1 < job.nblocks) {
        ib_finish_run(job, (long long)b1 * 256 + threadIdx.x);
        return;
    }
    long long bid = (long long)b1 - job.nblocks;
    const int bt = (int)(bid % a.n_btiles);
    bid /= a.n_btiles;
    const int ftile = (int)(bid % a.n_ftiles);
    bid /= a.n_ftiles;
    const int ttile = (int)(bid % a.n_ttiles);
    const long long heap = bid / a.n_ttiles;
    const int b0 = bt * kTrBeams, f0 = ftile * a.ft, t0 = ttile * a.tt;
    const int bv = min(kTrBeams, a.B - b0), fv = min(a.ft, a.Fo - f0), tv = min(a.tt, a.ohs - t0);
    const long long tg0 = heap * a.ohs + t0;

    // 16-byte path: beams, channels, B and Fo multiples of 16 (the tile's
    // beams and channels then start 16-aligned, and so do the rows).
    const bool wide = ((bv | fv | a.B | a.Fo) & 15) == 0;
    if (wide) {
        // Load item: 16 beams x 4 channels as four 16-byte loads (beam
        // sixteens fastest, so a wave reads whole 128-byte rows); four 4 x 4
        // byte transposes with v_perm_b32; sixteen words of 4 channels to
        // LDS.
        const int nb16 = bv >> 4, nfg = fv >> 2;
        const int nitems = nb16 * tv * nfg;
        const size_t fstep = (size_t)a.To * a.B;
        for (int i = threadIdx.x; i < nitems; i += 256) {
            const int bs = i % nb16;
            const int rest = i / nb16;
            const int t = rest % tv;
            const int fg = rest / tv;
            const int8_t* src = a.src + ((size_t)(f0 + 4 * fg) * a.To + tg0 + t) * a.B + b0 + 16 * bs;
            const uint4 q0 = *reinterpret_cast<const uint4*>(src);
            const uint4 q1 = *reinterpret_cast<const uint4*>(src + fstep);
            const uint4 q2 = *reinterpret_cast<const uint4*>(src + 2 * fstep);
            const uint4 q3 = *reinterpret_cast<const uint4*>(src + 3 * fstep);
            uint32_t* l = reinterpret_cast<uint32_t*>(tr_lds) + (size_t)t * a.tp + (size_t)(16 * bs) * a.bp + fg;
            auto quad = [&](uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3, uint32_t* lq) {
                const uint32_t t0 = __builtin_amdgcn_perm(r1, r0, 0x05010400u);
                const uint32_t t1 = __builtin_amdgcn_perm(r1, r0, 0x07030602u);
                const uint32_t t2 = __builtin_amdgcn_perm(r3, r2, 0x05010400u);
                const uint32_t t3 = __builtin_amdgcn_perm(r3, r2, 0x07030602u);
                lq[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
                lq[a.bp] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
                lq[2 * a.bp] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
                lq[3 * a.bp] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
            };
            quad(q0.x, q1.x, q2.x, q3.x, l);
            quad(q0.y, q1.y, q2.y, q3.y, l + 4 * a.bp);
            quad(q0.z, q1.z, q2.z, q3.z, l + 8 * a.bp);
            quad(q0.w, q1.w, q2.w, q3.w, l + 12 * a.bp);
        }
    } else if ((fv & 3) == 0 && (f0 & 3) == 0) {
        // Load 4 channels x 4 beams as four 4-byte words (beam quads
        // fastest, so a wave reads whole 128-byte rows), transpose the 4 x 4
        // bytes with v_perm_b32 and write four words of 4 channels.
        const int nbq = bv >> 2, nfg = fv >> 2;
        const int nitems = nbq * tv * nfg;
        for (int i = threadIdx.x; i < nitems; i += 256) {
            const int bq = i % nbq;
            const int rest = i / nbq;
            const int t = rest % tv;
            const int fg = rest / tv;
            const int8_t* src = a.src + ((size_t)(f0 + 4 * fg) * a.To + tg0 + t) * a.B + b0 + 4 * bq;
            const size_t fstep = (size_t)a.To * a.B;
            const uint32_t r0 = *reinterpret_cast<const uint32_t*>(src);
            const uint32_t r1 = *reinterpret_cast<const uint32_t*>(src + fstep);
            const uint32_t r2 = *reinterpret_cast<const uint32_t*>(src + 2 * fstep);
            const uint32_t r3 = *reinterpret_cast<const uint32_t*>(src + 3 * fstep);
            const uint32_t t0 = __builtin_amdgcn_perm(r1, r0, 0x05010400u);   // r0.0 r1.0 r0.1 r1.1
            const uint32_t t1 = __builtin_amdgcn_perm(r1, r0, 0x07030602u);   // r0.2 r1.2 r0.3 r1.3
            const uint32_t t2 = __builtin_amdgcn_perm(r3, r2, 0x05010400u);   // r2.0 r3.0 r2.1 r3.1
            const uint32_t t3 = __builtin_amdgcn_perm(r3, r2, 0x07030602u);   // r2.2 r3.2 r2.3 r3.3
            uint32_t* l = reinterpret_cast<uint32_t*>(tr_lds) + (size_t)t * a.tp + (size_t)(4 * bq) * a.bp + fg;
            l[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);               // beam 0: channels 0..3
            l[a.bp] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);            // beam 1
            l[2 * a.bp] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);        // beam 2
            l[3 * a.bp] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);        // beam 3
        }
    } else {
    // Load: (piece of 8 beams, t, f), pieces fastest.
    const int nb8 = bv >> 3;
    const int nload = fv * tv * nb8;
    for (int i = threadIdx.x; i < nload; i += 256) {
        const int pb = i % nb8;
        const int rest = i / nb8;
        const int t = rest % tv;
        const int f = rest / tv;
        const uint2 v = *reinterpret_cast<const uint2*>(a.src + ((size_t)(f0 + f) * a.To + tg0 + t) * a.B + b0 +
                                                        8 * pb);
        int8_t* l = tr_lds + ((size_t)t * a.tp + (size_t)(8 * pb) * a.bp) * 4 + f;
        const size_t bstep = (size_t)a.bp * 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            l[k * bstep] = (int8_t)(v.x >> (8 * k));
            l[(k + 4) * bstep] = (int8_t)(v.y >> (8 * k));
        }
    }
    }
    __syncthreads();

    // Store: (word of 4 channels, t, b), words fastest; a beam's tv rows of
    // Fo channels are contiguous in TBTF when the tile holds every channel.
    int8_t* out = a.dst + (((size_t)heap * a.B + b0) * a.ohs + t0) * a.Fo + f0;
    const size_t brow = (size_t)a.ohs * a.Fo;
    if (wide) {
        // Store item: 16 channels of one (beam, sample) as one 16-byte store
        // (channel sixteens fastest, then samples: a beam's tv rows are one
        // contiguous run).
        const int n16 = fv >> 4;
        const int nstore = bv * tv * n16;
        for (int i = threadIdx.x; i < nstore; i += 256) {
            const int w = i % n16;
            const int rest = i / n16;
            const int t = rest % tv;
            const int b = rest / tv;
            const uint32_t* l = reinterpret_cast<const uint32_t*>(tr_lds) + (size_t)t * a.tp + (size_t)b * a.bp + 4 * w;
            uint4 v;
            v.x = l[0];
            v.y = l[1];
            v.z = l[2];
            v.w = l[3];
            *reinterpret_cast<uint4*>(out + b * brow + (size_t)t * a.Fo + 16 * w) = v;
        }
    } else if ((a.Fo & 3) == 0) {
        const int nw = fv >> 2;
        const int nstore = bv * tv * nw;
        for (int i = threadIdx.x; i < nstore; i += 256) {
            const int w = i % nw;
            const int rest = i / nw;
            const int t = rest % tv;
            const int b = rest / tv;
            const uint32_t word =
                *reinterpret_cast<const uint32_t*>(tr_lds + ((size_t)t * a.tp + (size_t)b * a.bp + w) * 4);
            *reinterpret_cast<uint32_t*>(out + b * brow + (size_t)t * a.Fo + 4 * w) = word;
        }
    } else {
        const int nstore = bv * tv * fv;
        for (int i = threadIdx.x; i < nstore; i += 256) {
            const int f = i % fv;
            const int rest = i / fv;
            const int t = rest % tv;
            const int b = rest / tv;
            out[b * brow + (size_t)t * a.Fo + f] = tr_lds[((size_t)t * a.tp + (size_t)b * a.bp) * 4 + f];
        }
    }
}

// Tile shape and grid for tbtf_k.
TbtfArgs tbtf_args(const int8_t* src, int8_t* dst, long long To, int B, int Fo, int ohs)
{
    TbtfArgs a{};
    a.src = src;
    a.dst = dst;
    a.To = To;
    a.B = B;
    a.Fo = Fo;
    a.ohs = ohs;
    a.ft = Fo < kTrMaxChans ? Fo : kTrMaxChans;
    a.bp = (a.ft + 3) / 4 + 1;
    if ((a.bp & 1) == 0) ++a.bp;
    a.tp = kTrBeams * a.bp + 16;
    // Samples per tile: the fewest that still give each beam one 128-byte
    // output run (tt * Fo bytes when the tile holds every channel), so the
    // grid has as many work-groups as that allows (2 samples at 64
    // channels measured 1.5 % faster over the CB window than 4).
    int tt = 1;
    if (a.ft == Fo)
        while (tt < 16 && tt * Fo < 128) tt <<= 1;
    while (tt > 1 && (size_t)tt * a.tp * 4 > (size_t)kTrLdsBudget) tt >>= 1;
    a.tt = tt < ohs ? tt : ohs;
    a.n_ttiles = (ohs + a.tt - 1) / a.tt;
    a.n_ftiles = (Fo + a.ft - 1) / a.ft;
    a.n_btiles = (B + kTrBeams - 1) / kTrBeams;
    return a;
}

hipError_t launch_tbtf(const TbtfArgs& a, hipStream_t s, const IbFinishJob& job = IbFinishJob{},
                       const WeightsJob& wj = WeightsJob{})
{
    const long long heaps = a.To / a.ohs;
    const long long nblocks = heaps * a.n_ttiles * a.n_ftiles * a.n_btiles + job.nblocks + wj.nblocks;
    if (nblocks <= 0) return hipSuccess;
    if (nblocks >= (1ll << 31)) return hipErrorInvalidConfiguration;
    const size_t shm = (size_t)a.tt * a.tp * 4;
    tbtf_k<<<dim3((unsigned)nblocks), 256, shm, s>>>(a, job, wj);
    return hipGetLastError();
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
    float* d_levels;  // device: offsets[Fo], scales[Fo], then the thresholds [Fo][kThr] (int64)
    void* staging;    // [Fo][To][B] int8 between the beamformer and tbtf_k
    size_t staging_bytes;
    void* ftpa;       // FTPA of fbf_cb_beamform_taftp's split-transpose path
    size_t ftpa_bytes;
    void* ibp;        // incoherent-beam partials of fbf_cb_ib_beamform_taftp
    size_t ibp_bytes;
    hipEvent_t t_before, t_after;   // recorded around the coherent kernel (fbf_cb_plan_set_timing)
};

static long long* cb_thr(const fbf_cb_plan* plan)
{
    return reinterpret_cast<long long*>(plan->d_levels + 2 * (plan->d.nchans / plan->d.fscrunch));
}

// Grows one of the plan's device buffers (not while the stream is being
// captured).
static int cb_grow(void** buf, size_t* have, size_t bytes, hipStream_t s, void** out)
{
    if (*have < bytes) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        FBF_HIP_TRY(hipStreamIsCapturing(s, &cs));
        FBF_REQUIRE(cs == hipStreamCaptureStatusNone, FBF_ERR_STATE,
                    "fbf_cb: the plan's buffers must be sized (fbf_cb_plan_reserve) before stream capture");
        // Earlier launches on any stream may still use the old buffer.
        if (*buf) {
            FBF_HIP_TRY(hipDeviceSynchronize());
            FBF_HIP_TRY(hipFree(*buf));
        }
        *buf = nullptr;
        *have = 0;
        FBF_HIP_TRY(hipMalloc(buf, bytes));
        *have = bytes;
    }
    *out = *buf;
    return FBF_OK;
}

static int cb_staging(fbf_cb_plan* plan, size_t bytes, hipStream_t s, void** out)
{
    return cb_grow(&plan->staging, &plan->staging_bytes, bytes, s, out);
}

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
    auto* p = new fbf_cb_plan{*desc, ohs, 0, 0, nullptr, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr};
    hipError_t e = hipMalloc(&p->d_levels, sizeof(float) * 2 * Fo + sizeof(long long) * kThr * Fo);
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
    // A beamform still running on any stream reads the levels and thresholds
    // being replaced (the library's streams do not wait for the null stream):
    // it finishes first.
    FBF_HIP_TRY(hipDeviceSynchronize());
    FBF_HIP_TRY(hipMemcpy(plan->d_levels, offsets, sizeof(float) * Fo, hipMemcpyHostToDevice));
    FBF_HIP_TRY(hipMemcpy(plan->d_levels + Fo, scales, sizeof(float) * Fo, hipMemcpyHostToDevice));
    cb_thresholds_k<<<Fo, 256, 0, nullptr>>>(plan->d_levels, plan->d_levels + Fo, plan->d.quant_mode,
                                             cb_thr(plan));
    FBF_HIP_TRY(hipGetLastError());
    FBF_HIP_TRY(hipStreamSynchronize(nullptr));
    return FBF_OK;
}

extern "C" int fbf_cb_plan_destroy(fbf_cb_plan* plan)
{
    if (!plan) return FBF_OK;
    hipError_t e = hipFree(plan->d_levels);
    if (plan->staging) {
        const hipError_t e2 = hipFree(plan->staging);
        if (e == hipSuccess) e = e2;
    }
    if (plan->ftpa) {
        const hipError_t e2 = hipFree(plan->ftpa);
        if (e == hipSuccess) e = e2;
    }
    if (plan->ibp) {
        const hipError_t e2 = hipFree(plan->ibp);
        if (e == hipSuccess) e = e2;
    }
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
    bool tr;         // transposed tiles (cb_tr_k): voltages as the A operand
    bool ws;         // transposed tiles with wave-private staging (cb_ws_k)
};
static const CbVariant kCbVariants[] = {
    {0, 0, false, false, false},   // auto
    {32, 4, false, false, false},  // 64 beams x 32 samples per wave
    {32, 2, false, false, false},  // 32 beams x 32 samples per wave
    {16, 8, false, false, false},  // 64 beams x 16 samples per wave
    {16, 4, false, false, false},  // 32 beams x 16 samples per wave
    {32, 1, true, false, false},   // LDS: 16 beams per wave, 64 per work-group
    {32, 2, true, false, false},   // LDS: 32 beams per wave, 128 per work-group
    {32, 4, true, false, false},   // LDS: 64 beams per wave, 256 per work-group
    {32, 2, true, true, false},    // transposed: 32 beams per wave
    {32, 4, true, true, false},    // transposed: 64 beams per wave
    {32, 4, true, true, true},     // transposed, wave-private staging: 64 beams per wave
    {32, 8, true, true, false},    // transposed: 128 beams per wave, two waves per SIMD
};
constexpr int kCbVariantLds = 5;
constexpr int kCbVariantTr2 = 8;
constexpr int kCbVariantTr4 = 9;
constexpr int kCbVariantWs = 10;
constexpr int kCbVariantTr8 = 11;

static bool cb_tr_applies(const fbf_cb_desc& d, int rt, int mode, long long T)
{
    return mode == MODE_INT8 && T % 16 == 0 && 2 * T / 32 < (1ll << 31) && T / d.tscrunch < (1ll << 31) &&
           cb_tr_supported(d.nantennas, d.tscrunch, d.fscrunch, rt);
}

static bool cb_ws_applies(const fbf_cb_desc& d, int mode, long long T)
{
    return mode == MODE_INT8 && T % 16 == 0 && 2 * T / 32 < (1ll << 31) &&
           cb_ws_supported(d.nantennas, d.tscrunch, d.fscrunch);
}

// Automatic choice: up to three 32-byte K steps (48 antennas) the
// wave-private kernel, which measured 20-85 % faster there (8 to 48
// antennas, tscrunch 4 to 64; one exception, 32 antennas at tscrunch 4, 8 %
// slower); else the transposed kernel with 128 beams per wave (fscrunch 1:
// 3.2 % faster than 64 beams per wave at 64 antennas, 64 channels, 512
// beams), else with 64, else with 32 beams per wave, else the LDS-staged
// kernel (which also serves int32 sums, power output and the geometries the
// transposed kernels do not).
static CbVariant cb_pick_variant(const fbf_cb_desc& d, int variant, int mode, long long T)
{
    if (variant <= 0) {
        const int ksteps = (2 * d.nantennas + 31) / 32;
        if (ksteps <= 3 && cb_ws_applies(d, mode, T)) return kCbVariants[kCbVariantWs];
        if (ksteps > 3 && cb_tr_applies(d, 8, mode, T)) return kCbVariants[kCbVariantTr8];
        if (cb_tr_applies(d, 4, mode, T)) return kCbVariants[kCbVariantTr4];
        if (cb_tr_applies(d, 2, mode, T)) return kCbVariants[kCbVariantTr2];
        return kCbVariants[kCbVariantLds];
    }
    const CbVariant v = kCbVariants[variant];
    if (v.ws && !cb_ws_applies(d, mode, T)) return cb_pick_variant(d, 0, mode, T);
    if (v.tr && !v.ws && !cb_tr_applies(d, v.rows, mode, T)) return kCbVariants[kCbVariantLds];
    return v;
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

// A TAFTP block read directly by the coherent beamformer: S samples per
// heap, Atot antennas, the beamformer's antennas from antenna a0.
struct TaftpView {
    int S, Atot, a0;
    int variant;   // the kernel variant to pick from (cb_taftp_variant)
    int* ibp;      // incoherent-beam partials to leave (cb_ib_fused), or null
    IbFinishJob ibj;   // their finish, run in the reorder's launch
    WeightsJob wj;     // weights generated in the reorder's launch (the next block's)
};

// Validates and launches one coherent-beam kernel over T samples of voltage
// columns that are `vstride` bytes apart, the beamformer's antennas starting
// `voff` bytes into each column; or, with `tv`, over the TAFTP block `ftpa`
// (the caller has checked cb_taftp_direct).
static int cb_launch(fbf_cb_plan* plan, int mode, const void* ftpa, long long T, int vstride, int voff,
                     const void* fba, size_t weights_bytes, void* out, size_t out_bytes, fbf_stream_t stream,
                     void* workspace = nullptr, size_t workspace_bytes = 0, const TaftpView* tv = nullptr)
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
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* kout = out;
    if (mode == MODE_INT8) {
        if (workspace) {
            FBF_REQUIRE(workspace_bytes >= expect && (reinterpret_cast<uintptr_t>(workspace) & 15) == 0,
                        FBF_ERR_SIZE, "fbf_cb: workspace too small or not 16-byte aligned");
            kout = workspace;
        } else {
            const int st = cb_staging(plan, expect, s, &kout);
            if (st) return st;
        }
    }

    const CbVariant var = cb_pick_variant(d, tv ? tv->variant : plan->variant, mode, T);
    // The timing events (fbf_cb_plan_set_timing), around the coherent kernel only.
    auto mark = [&](hipEvent_t ev) { return ev ? hipEventRecord(ev, s) : hipSuccess; };
    FBF_REQUIRE(!tv || (var.tr && !var.ws && var.rows != 2 && d.fscrunch == 1), FBF_ERR_STATE,
                "fbf_cb: the TAFTP input needs the transposed kernel with 64 or 128 beams per wave, fscrunch 1");
    if (var.ws) {
        CbWsArgs wa{};
        CbArgs& a = wa.p;
        a.ftpa = static_cast<const int8_t*>(ftpa);
        a.fba = static_cast<const int8_t*>(fba);
        a.out = kout;
        a.offsets = plan->d_levels;
        a.scales = plan->d_levels + Fo;
        a.thr = cb_thr(plan);
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
        const int W = 2 * d.tscrunch;
        const int spt = W >= 32 ? 1 : 32 / W;
        wa.tps = W >= 32 ? W / 32 : 1;
        wa.ntp = (cb_ws_nslot(W) / spt) * wa.tps;
        wa.ntiles = (int)(2 * T / 32);
        wa.nbg = (d.nbeams + 63) / 64;
        const int ksteps = (2 * d.nantennas + 31) / 32;
        hipError_t e = mark(plan->t_before);
        if (e == hipSuccess) e = launch_ws(wa, W, ksteps, s);
        if (e == hipSuccess) e = mark(plan->t_after);
        if (e != hipSuccess) return hip_status(e, "cb_ws_k launch");
        e = launch_tbtf(tbtf_args(static_cast<const int8_t*>(kout), static_cast<int8_t*>(out), To, d.nbeams, Fo,
                                  plan->ohs),
                        s);
        if (e != hipSuccess) return hip_status(e, "tbtf_k launch");
        return FBF_OK;
    }
    if (var.tr) {
        CbTrArgs ta{};
        CbArgs& a = ta.p;
        a.ftpa = static_cast<const int8_t*>(ftpa);
        a.fba = static_cast<const int8_t*>(fba);
        a.out = kout;
        a.offsets = plan->d_levels;
        a.scales = plan->d_levels + Fo;
        a.thr = cb_thr(plan);
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
        if (tv) {
            a.S = tv->S;
            a.Atot = tv->Atot;
            a.a0 = tv->a0;
            ta.hstride = (size_t)4 * tv->Atot * d.nchans * tv->S;
            ta.lgtph = __builtin_ctz((unsigned)(tv->S / 16));
            ta.ibp = tv->ibp;
        }
        const int rt = var.rows;
        const int W = 2 * d.tscrunch;
        const int spt = W >= 32 ? 1 : 32 / W;
        ta.tps = W >= 32 ? W / 32 : 1;
        const int nslot = cb_tr_nslot(rt, W);
        ta.ntp = (nslot / spt) * ta.tps;
        ta.ntiles = (int)(2 * T / 32);
        const int passes = (ta.ntiles + ta.ntp - 1) / ta.ntp;
        const int beams_per_wg = 16 * rt * (4 / d.fscrunch);
        ta.n_bgroups = (d.nbeams + beams_per_wg - 1) / beams_per_wg;
        const int ksteps = (2 * d.nantennas + 31) / 32;
        hipError_t e = mark(plan->t_before);
        if (e == hipSuccess)
            e = rt == 2   ? launch_tr<2>(ta, d.fscrunch, W, ksteps, passes, s)
                : rt == 4 ? launch_tr<4>(ta, d.fscrunch, W, ksteps, passes, s)
                          : launch_tr<8>(ta, d.fscrunch, W, ksteps, passes, s);
        if (e == hipSuccess) e = mark(plan->t_after);
        if (e != hipSuccess) return hip_status(e, "cb_tr_k launch");
        e = launch_tbtf(tbtf_args(static_cast<const int8_t*>(kout), static_cast<int8_t*>(out), To, d.nbeams, Fo,
                                  plan->ohs),
                        s, tv ? tv->ibj : IbFinishJob{}, tv ? tv->wj : WeightsJob{});
        if (e != hipSuccess) return hip_status(e, "tbtf_k launch");
        return FBF_OK;
    }
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
    CbArgs a{};
    a.ftpa = static_cast<const int8_t*>(ftpa);
    a.fba = static_cast<const int8_t*>(fba);
    a.out = kout;
    a.offsets = plan->d_levels;
    a.scales = plan->d_levels + Fo;
    a.thr = cb_thr(plan);
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
    dim3 grid((unsigned)nblocks);
    hipError_t e = mark(plan->t_before);
    if (e != hipSuccess) return hip_status(e, "hipEventRecord");
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
    if (e == hipSuccess) e = mark(plan->t_after);
    if (e != hipSuccess) return hip_status(e, "cb_mfma_k launch");
    if (mode == MODE_INT8) {
        e = launch_tbtf(tbtf_args(static_cast<const int8_t*>(kout), static_cast<int8_t*>(out), To, d.nbeams, Fo,
                                  plan->ohs),
                        s);
        if (e != hipSuccess) return hip_status(e, "tbtf_k launch");
    }
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

// Whether fbf_cb_beamform_taftp stages the TAFTP block straight into the
// transposed kernel: the plan's kernel is cb_tr_k with 64 or 128 beams per
// wave and fscrunch 1, a heap holds a power-of-two number of samples, at least one tile
// (16 samples), and a heap group's bytes fit a 31-bit offset.
// The variant fbf_cb_beamform_taftp picks from: the plan's; when automatic,
// 128 beams per wave above three K steps where that applies (the automatic
// FTPA pick as well).  That reads each staged tile once per 128 beams, which
// halves the transposed LDS reads per MFMA against 64 beams per wave.
static int cb_taftp_variant(const fbf_cb_plan* plan, long long T)
{
    const fbf_cb_desc& d = plan->d;
    if (plan->variant > 0) return plan->variant;
    if ((2 * d.nantennas + 31) / 32 > 3 && d.fscrunch == 1 && cb_tr_applies(d, 8, MODE_INT8, T))
        return kCbVariantTr8;
    return 0;
}

static bool cb_taftp_direct(const fbf_cb_plan* plan, long long T, int nantennas_total, int S)
{
    const fbf_cb_desc& d = plan->d;
    const CbVariant var = cb_pick_variant(d, cb_taftp_variant(plan, T), MODE_INT8, T);
    return var.tr && !var.ws && var.rows != 2 && d.fscrunch == 1 && S >= 16 && (S & (S - 1)) == 0 &&
           (long long)4 * nantennas_total * d.nchans * S < (1ll << 31) &&
           (long long)4 * nantennas_total * d.nchans * T < (1ll << 32);   // 32-bit tile offsets
}

extern "C" int fbf_cb_beamform_taftp(fbf_cb_plan* plan, const void* taftp, size_t taftp_bytes, int nantennas_total,
                                     int start_antenna, int nsamples_per_heap, const void* fba_weights,
                                     size_t weights_bytes, void* tbtf, size_t tbtf_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb_beamform_taftp: null plan");
    FBF_REQUIRE(taftp && fba_weights && tbtf, FBF_ERR_ARG, "fbf_cb_beamform_taftp: null buffer");
    const fbf_cb_desc& d = plan->d;
    FBF_REQUIRE(nantennas_total > 0 && start_antenna >= 0 && start_antenna + d.nantennas <= nantennas_total,
                FBF_ERR_ARG, "fbf_cb_beamform_taftp: the plan's antennas lie outside the block's");
    FBF_REQUIRE(nsamples_per_heap > 0, FBF_ERR_ARG, "fbf_cb_beamform_taftp: nsamples_per_heap must be positive");
    const size_t hg_bytes = (size_t)4 * nantennas_total * d.nchans * nsamples_per_heap;
    FBF_REQUIRE(taftp_bytes % hg_bytes == 0, FBF_ERR_SIZE,
                "fbf_cb_beamform_taftp: input is not a whole number of heap groups");
    const long long T = (long long)(taftp_bytes / hg_bytes) * nsamples_per_heap;
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(taftp) & 15) == 0, FBF_ERR_ARG,
                "fbf_cb_beamform_taftp: buffers must be 16-byte aligned");
    if (T == 0 || cb_taftp_direct(plan, T, nantennas_total, nsamples_per_heap)) {
        const TaftpView tv{nsamples_per_heap, nantennas_total, start_antenna, cb_taftp_variant(plan, T), nullptr,
                           IbFinishJob{}, WeightsJob{}};
        return cb_launch(plan, MODE_INT8, taftp, T, 2 * d.nantennas, 0, fba_weights, weights_bytes, tbtf,
                         tbtf_bytes, stream, nullptr, 0, T ? &tv : nullptr);
    }
    // Otherwise the plan's antennas are split-transposed into the plan's FTPA
    // buffer for the FTPA kernels.  Sizes are checked before any launch.
    const size_t fbytes = (size_t)4 * d.nantennas * d.nchans * T;
    FBF_REQUIRE(weights_bytes == (size_t)2 * d.nchans * d.nbeams * d.nantennas, FBF_ERR_SIZE,
                "fbf_cb_beamform_taftp: weights size " + std::to_string(weights_bytes) + " != expected " +
                    std::to_string((size_t)2 * d.nchans * d.nbeams * d.nantennas));
    size_t expect = 0;
    int st = fbf_cb_output_bytes(plan, fbytes, &expect);
    if (st) return st;
    FBF_REQUIRE(tbtf_bytes == expect && (T / d.tscrunch) % plan->ohs == 0, FBF_ERR_SIZE,
                "fbf_cb_beamform_taftp: output size " + std::to_string(tbtf_bytes) + " != expected " +
                    std::to_string(expect) + ", or not whole output heaps");
    void* ftpa = nullptr;
    st = cb_grow(&plan->ftpa, &plan->ftpa_bytes, fbytes, static_cast<hipStream_t>(stream), &ftpa);
    if (st) return st;
    st = fbf_split_transpose(taftp, taftp_bytes, ftpa, fbytes, nantennas_total, d.nantennas, start_antenna,
                             d.nchans, nsamples_per_heap, stream);
    if (st) return st;
    return cb_launch(plan, MODE_INT8, ftpa, T, 2 * d.nantennas, 0, fba_weights, weights_bytes, tbtf, tbtf_bytes,
                     stream);
}

// Whether fbf_cb_ib_beamform_taftp forms the incoherent beam inside the
// coherent beamformer: the direct TAFTP path with 128 beams per wave over
// all the block's antennas, and an incoherent-beam plan of the same block
// (antennas, channels, heap size).
static bool cb_ib_fused(const fbf_cb_plan* plan, const IbFuseParams& ibp, long long T, int S)
{
    const fbf_cb_desc& d = plan->d;
    const CbVariant var = cb_pick_variant(d, cb_taftp_variant(plan, T), MODE_INT8, T);
    return cb_taftp_direct(plan, T, d.nantennas, S) && var.rows == 8 && ibp.nantennas == d.nantennas &&
           ibp.nchans == d.nchans && ibp.nsamples_per_heap == S;
}

// Both beams of a TAFTP block in one pass (or one after the other where the
// plans do not allow it); `wjob`, when it has work-groups, generates weights
// in the reorder's launch, else `next` (a fbf_weights_generate call) runs
// after the beams.
struct NextWeights {
    fbf_weights_plan* plan;
    const void* delays;
    size_t delays_bytes;
    double epoch;
    void* weights;
    size_t weights_bytes;
};

static int cb_ib_run(fbf_cb_plan* plan, fbf_ib_plan* ib_plan, const void* taftp, size_t taftp_bytes,
                     int nsamples_per_heap, const void* fba_weights, size_t weights_bytes, void* tbtf,
                     size_t tbtf_bytes, void* tf, size_t tf_bytes, fbf_stream_t stream, const NextWeights* next)
{
    FBF_REQUIRE(plan && ib_plan, FBF_ERR_ARG, "fbf_cb_ib_beamform_taftp: null plan");
    FBF_REQUIRE(taftp && fba_weights && tbtf && tf, FBF_ERR_ARG, "fbf_cb_ib_beamform_taftp: null buffer");
    IbFuseParams ibp;
    int st = ib_fuse_params(ib_plan, &ibp);
    if (st) return st;
    const fbf_cb_desc& d = plan->d;
    FBF_REQUIRE(nsamples_per_heap > 0, FBF_ERR_ARG, "fbf_cb_ib_beamform_taftp: nsamples_per_heap must be positive");
    const size_t hg_bytes = (size_t)4 * d.nantennas * d.nchans * nsamples_per_heap;
    FBF_REQUIRE(taftp_bytes % hg_bytes == 0, FBF_ERR_SIZE,
                "fbf_cb_ib_beamform_taftp: input is not a whole number of heap groups of the plan's antennas");
    const long long T = (long long)(taftp_bytes / hg_bytes) * nsamples_per_heap;
    WeightsJob wj{};
    if (next) {
        // Checked before anything is launched.
        st = weights_launch_params(next->plan, next->delays, next->delays_bytes, next->epoch, next->weights,
                                   next->weights_bytes, &wj.w);
        if (st) return st;
        FBF_REQUIRE(next->weights_bytes == weights_bytes, FBF_ERR_SIZE,
                    "fbf_cb_ib_beamform_taftp_next: the next weights differ in size from the block's");
        FBF_REQUIRE(wj.w.F == d.nchans && (long long)wj.w.nd == (long long)d.nantennas * d.nbeams, FBF_ERR_ARG,
                    "fbf_cb_ib_beamform_taftp_next: the weights plan's geometry differs from the coherent plan's");
        FBF_REQUIRE((unsigned long long)wj.w.gx * wj.w.gy < (1ull << 30), FBF_ERR_UNSUPPORTED,
                    "fbf_cb_ib_beamform_taftp_next: too many weights work-groups");
        wj.nblocks = (int)(wj.w.gx * wj.w.gy);
    }
    if (T == 0 || !cb_ib_fused(plan, ibp, T, nsamples_per_heap)) {
        // The two beamformers one after the other, then the next weights.
        st = fbf_ib_beamform(ib_plan, taftp, taftp_bytes, tf, tf_bytes, stream);
        if (st) return st;
        st = fbf_cb_beamform_taftp(plan, taftp, taftp_bytes, d.nantennas, 0, nsamples_per_heap, fba_weights,
                                   weights_bytes, tbtf, tbtf_bytes, stream);
        if (st || !next) return st;
        return fbf_weights_generate(next->plan, next->delays, next->delays_bytes, next->epoch, next->weights,
                                    next->weights_bytes, nullptr, stream);
    }
    size_t tf_expect = 0;
    st = fbf_ib_output_bytes(ib_plan, taftp_bytes, &tf_expect);
    if (st) return st;
    FBF_REQUIRE(tf_bytes == tf_expect, FBF_ERR_SIZE,
                "fbf_cb_ib_beamform_taftp: incoherent-beam output size " + std::to_string(tf_bytes) +
                    " != expected " + std::to_string(tf_expect));
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(taftp) & 15) == 0, FBF_ERR_ARG,
                "fbf_cb_ib_beamform_taftp: buffers must be 16-byte aligned");
    void* part = nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    st = cb_grow(&plan->ibp, &plan->ibp_bytes, (size_t)d.nchans * T * 4 * sizeof(int), s, &part);
    if (st) return st;
    IbFinishJob ibj{};
    st = ib_finish_job(ib_plan, static_cast<const int*>(part), T, tf, tf_bytes, &ibj);
    if (st) return st;
    const TaftpView tv{nsamples_per_heap, d.nantennas, 0, cb_taftp_variant(plan, T), static_cast<int*>(part), ibj,
                       wj};
    return cb_launch(plan, MODE_INT8, taftp, T, 2 * d.nantennas, 0, fba_weights, weights_bytes, tbtf, tbtf_bytes,
                     stream, nullptr, 0, &tv);
}

extern "C" int fbf_cb_ib_beamform_taftp(fbf_cb_plan* plan, fbf_ib_plan* ib_plan, const void* taftp,
                                        size_t taftp_bytes, int nsamples_per_heap, const void* fba_weights,
                                        size_t weights_bytes, void* tbtf, size_t tbtf_bytes, void* tf,
                                        size_t tf_bytes, fbf_stream_t stream)
{
    return cb_ib_run(plan, ib_plan, taftp, taftp_bytes, nsamples_per_heap, fba_weights, weights_bytes, tbtf,
                     tbtf_bytes, tf, tf_bytes, stream, nullptr);
}

extern "C" int fbf_cb_ib_beamform_taftp_next(fbf_cb_plan* plan, fbf_ib_plan* ib_plan, const void* taftp,
                                             size_t taftp_bytes, int nsamples_per_heap, const void* fba_weights,
                                             size_t weights_bytes, void* tbtf, size_t tbtf_bytes, void* tf,
                                             size_t tf_bytes, fbf_weights_plan* weights_plan, const void* delays,
                                             size_t delays_bytes, double next_epoch, void* next_weights,
                                             size_t next_weights_bytes, fbf_stream_t stream)
{
    const NextWeights next{weights_plan, delays, delays_bytes, next_epoch, next_weights, next_weights_bytes};
    return cb_ib_run(plan, ib_plan, taftp, taftp_bytes, nsamples_per_heap, fba_weights, weights_bytes, tbtf,
                     tbtf_bytes, tf, tf_bytes, stream, &next);
}

extern "C" int fbf_cb_ib_fused(const fbf_cb_plan* plan, const fbf_ib_plan* ib_plan, int nsamples_per_heap,
                               long long nsamples, int* fused)
{
    FBF_REQUIRE(plan && ib_plan && fused, FBF_ERR_ARG, "fbf_cb_ib_fused: null argument");
    IbFuseParams ibp;
    const int st = ib_fuse_params(ib_plan, &ibp);
    if (st) return st;
    *fused = nsamples > 0 && nsamples % plan->d.tscrunch == 0 && cb_ib_fused(plan, ibp, nsamples, nsamples_per_heap);
    return FBF_OK;
}

extern "C" int fbf_cb_taftp_direct(const fbf_cb_plan* plan, int nantennas_total, int nsamples_per_heap,
                                   long long nsamples, int* direct)
{
    FBF_REQUIRE(plan && direct, FBF_ERR_ARG, "fbf_cb_taftp_direct: null argument");
    *direct = nsamples > 0 && nsamples % plan->d.tscrunch == 0 &&
              cb_taftp_direct(plan, nsamples, nantennas_total, nsamples_per_heap);
    return FBF_OK;
}

extern "C" int fbf_cb_plan_set_timing(fbf_cb_plan* plan, fbf_event_t before, fbf_event_t after)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb_plan_set_timing: null plan");
    FBF_REQUIRE((before == nullptr) == (after == nullptr), FBF_ERR_ARG,
                "fbf_cb_plan_set_timing: give both events or neither");
    plan->t_before = static_cast<hipEvent_t>(before);
    plan->t_after = static_cast<hipEvent_t>(after);
    return FBF_OK;
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

extern "C" int fbf_cb_beamform_ws(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes, const void* fba_weights,
                                  size_t weights_bytes, void* tbtf, size_t tbtf_bytes, void* workspace,
                                  size_t workspace_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb: null plan");
    FBF_REQUIRE(workspace, FBF_ERR_ARG, "fbf_cb_beamform_ws: null workspace");
    long long T = 0;
    int st = cb_nsamples(plan, ftpa_bytes, &T);
    if (st) return st;
    return cb_launch(plan, MODE_INT8, ftpa, T, 2 * plan->d.nantennas, 0, fba_weights, weights_bytes, tbtf,
                     tbtf_bytes, stream, workspace, workspace_bytes);
}

extern "C" int fbf_cb_plan_reserve(fbf_cb_plan* plan, long long nsamples)
{
    FBF_REQUIRE(plan && nsamples >= 0, FBF_ERR_ARG, "fbf_cb_plan_reserve: bad argument");
    FBF_REQUIRE(nsamples % plan->d.tscrunch == 0, FBF_ERR_SIZE, "fbf_cb_plan_reserve: nsamples % tscrunch != 0");
    const size_t bytes = (size_t)(nsamples / plan->d.tscrunch) * plan->d.nbeams * (plan->d.nchans / plan->d.fscrunch);
    if (bytes == 0) return FBF_OK;
    void* p = nullptr;
    const int st = cb_staging(plan, bytes, nullptr, &p);
    if (st) return st;
    // The incoherent-beam partials of fbf_cb_ib_beamform_taftp (16 bytes per
    // channel and sample).
    return cb_grow(&plan->ibp, &plan->ibp_bytes, (size_t)plan->d.nchans * nsamples * 4 * sizeof(int), nullptr, &p);
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
        if (kCbVariants[var].lds && !kCbVariants[var].tr)
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
