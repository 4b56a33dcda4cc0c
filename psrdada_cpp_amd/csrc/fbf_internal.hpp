// Internal helpers shared by the HIP kernels and the C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/fbfuse_hip.h"

namespace fbf {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
void clear_error();

// Returns FBF_ERR_HIP (and records the HIP error string) when `e` failed.
int hip_status(hipError_t e, const char* what);

}  // namespace fbf

#define FBF_HIP_TRY(expr)                                              \
    do {                                                               \
        hipError_t fbf_e_ = (expr);                                    \
        if (fbf_e_ != hipSuccess) return ::fbf::hip_status(fbf_e_, #expr); \
    } while (0)

#define FBF_REQUIRE(cond, code, msg)                                   \
    do {                                                               \
        if (!(cond)) return ::fbf::set_error((code), (msg));           \
    } while (0)

// Per-launch check: hipGetLastError after a kernel launch.
#define FBF_LAUNCH_CHECK(name)                                         \
    do {                                                               \
        hipError_t fbf_e_ = hipGetLastError();                         \
        if (fbf_e_ != hipSuccess) return ::fbf::hip_status(fbf_e_, name); \
    } while (0)

namespace fbf {

// What the fused split-transpose + incoherent-beam path needs from an
// incoherent-beam plan (incoherent_beamformer.hip).
struct IbFuseParams {
    const float* offsets;   // device, [nchans / fscrunch]
    const float* scales;
    int nantennas, nchans, nsamples_per_heap, tscrunch, fscrunch, quant;
};
int ib_fuse_params(const fbf_ib_plan* plan, IbFuseParams* out);

// Weights generation (weights.hip), shared by weights_k and by launches that
// generate the weights beside other work.  A thread takes one (beam,
// antenna) di over the run of kWChans channels starting at f0.
constexpr int kWChans = 8;

struct WeightsLaunch {
    const float2* delays;   // [B][A] (rate, offset)
    const double* freqs;    // device, [F]
    double epoch;
    int F, nd;              // channels; beams x antennas
    char2* weights;         // [F][B][A]
    unsigned gx, gy;        // work-groups over (beam, antenna) pairs and channel runs
};
// Checks a generate call's arguments against its plan and fills the launch.
int weights_launch_params(const fbf_weights_plan* plan, const void* delays, size_t delays_bytes, double epoch,
                          void* weights, size_t weights_bytes, WeightsLaunch* out);

// phase = (rate * t + offset) * f in FP64 (contraction off, the oracle's
// order), the whole-cycle part removed, then FP32 sincospif of the reduced
// phase and both components rounded to the nearest integer, ties to even.
__device__ __forceinline__ void weights_run(const WeightsLaunch& w, float* phase_out, int di, int f0)
{
#pragma clang fp contract(off)
    const float2 d = w.delays[di];
    double tau = (double)d.x * w.epoch;
    tau = tau + (double)d.y;
    const int f1 = f0 + kWChans < w.F ? f0 + kWChans : w.F;
    for (int f = f0; f < f1; ++f) {
        const double phase = tau * w.freqs[f];
        const double frac = phase - rint(phase);
        const float x = (float)frac;
        float s, c;
        sincospif(2.0f * x, &s, &c);
        const float re = 127.0f * c;
        const float im = 127.0f * s;
        char2 v;
        v.x = (signed char)(int)rintf(re);
        v.y = (signed char)(int)rintf(im);
        const size_t idx = (size_t)f * w.nd + di;
        w.weights[idx] = v;
        if (phase_out) phase_out[idx] = x;
    }
}

// Number of XCDs on MI355X; work-groups are dispatched round robin over them.
constexpr int kNumXcd = 8;

// Maps a hardware block id to a logical tile id so that logical neighbours
// (tiles that share input through L2) run on the same XCD.  Bijective on
// [0, nblocks).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int per = nblocks / kNumXcd;
    if (bid >= per * kNumXcd) return bid;
    return (bid % kNumXcd) * per + bid / kNumXcd;
}

// Quantiser shared by every output path: (P - offset) / scale in IEEE double,
// round half away from zero (mode 0) or truncate (mode 1), saturate to int8.
__device__ __forceinline__ int8_t quantise(double power, float offset, float scale, int mode)
{
    double v = (power - (double)offset) / (double)scale;
    double q = mode ? trunc(v) : round(v);
    q = fmax(q, -128.0);  // fmax maps NaN to -128
    q = fmin(q, 127.0);
    return (int8_t)(int)q;
}

}  // namespace fbf
