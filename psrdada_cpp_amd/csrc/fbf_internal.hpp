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
