// Delay model -> int8 beamforming weights.
//
// Behaviour follows WeightsManager.cu:14-123 (generate_weights_k and
// WeightsManager::weights): for delay model (rate, offset) of (beam, antenna)
// and channel frequency f,
//   phase = (t * rate + offset) * f          [cycles]
//   w     = round(127 * exp(i * 2 pi * phase))   -> char2 (re = cos, im = sin)
// written [F][B][A] (the reference's [t][F][B][A] with one time step).
//
// MI355X design: the phase argument is formed in FP64 -- the only FP64 step --
// because f * tau reaches thousands of cycles at GHz x microseconds and FP32
// would leave ~1e-3 rad of error.  The whole-cycle part is removed in FP64
// (phase - rint(phase), exact), the reduced argument in [-1/2, 1/2] goes to
// FP32 sincospif (HIP's device math library), then rounds to the nearest
// integer, ties to even, as __float2int_rn does in WeightsManager.cu:70-71.
// A thread takes one (beam, antenna) over a run of kWChans channels: the
// delay model is read and tau formed once, and each channel costs one FP64
// product, the reduction, sincospif and the rounding.  Antennas are fastest
// across lanes, so the delay reads and each channel's weight writes are
// coalesced; the grid's y dimension runs over the channel runs.
#include "fbf_internal.hpp"

namespace fbf {
namespace {

__global__ __launch_bounds__(256) void weights_k(WeightsLaunch w, float* __restrict__ phase_out)
{
    const int di = blockIdx.x * blockDim.x + threadIdx.x;   // b * A + a
    if (di >= w.nd) return;
    weights_run(w, phase_out, di, blockIdx.y * kWChans);
}

// Time steps of one delay model (WeightsManager.cu:14-79's tstart, tstep,
// ntsteps): step k is blockIdx.z, at epoch tstart + k * tstep formed in
// FP64 (product, then sum), written at weights + k * F * B * A.
__global__ __launch_bounds__(256) void weights_steps_k(WeightsLaunch w, double tstart, double tstep)
{
#pragma clang fp contract(off)
    const int di = blockIdx.x * blockDim.x + threadIdx.x;
    if (di >= w.nd) return;
    const int k = blockIdx.z;
    WeightsLaunch wk = w;
    const double dt = (double)k * tstep;
    wk.epoch = tstart + dt;
    wk.weights = w.weights + (size_t)k * w.F * w.nd;
    weights_run(wk, nullptr, di, blockIdx.y * kWChans);
}

}  // namespace
}  // namespace fbf

using namespace fbf;

struct fbf_weights_plan {
    int A, B, F;
    double* d_freqs;
};

extern "C" int fbf_weights_plan_create(int nantennas, int nbeams, int nchans, const double* chan_freqs,
                                       fbf_weights_plan** plan)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_weights_plan_create: null plan pointer");
    *plan = nullptr;
    FBF_REQUIRE(nantennas > 0 && nbeams > 0 && nchans > 0 && chan_freqs, FBF_ERR_ARG,
                "fbf_weights_plan_create: bad geometry or null frequencies");
    auto* p = new fbf_weights_plan{nantennas, nbeams, nchans, nullptr};
    hipError_t e = hipMalloc(&p->d_freqs, sizeof(double) * nchans);
    if (e == hipSuccess) e = hipMemcpy(p->d_freqs, chan_freqs, sizeof(double) * nchans, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->d_freqs) (void)hipFree(p->d_freqs);
        delete p;
        return hip_status(e, "fbf_weights_plan_create");
    }
    *plan = p;
    return FBF_OK;
}

extern "C" int fbf_weights_plan_destroy(fbf_weights_plan* plan)
{
    if (!plan) return FBF_OK;
    hipError_t e = hipFree(plan->d_freqs);
    delete plan;
    if (e != hipSuccess) return hip_status(e, "hipFree(freqs)");
    return FBF_OK;
}

int fbf::weights_launch_params(const fbf_weights_plan* plan, const void* delays, size_t delays_bytes, double epoch,
                               void* weights, size_t weights_bytes, WeightsLaunch* out)
{
    FBF_REQUIRE(plan && delays && weights, FBF_ERR_ARG, "fbf_weights_generate: null argument");
    const size_t nd = (size_t)plan->A * plan->B;
    FBF_REQUIRE(delays_bytes == nd * sizeof(float2), FBF_ERR_SIZE,
                "fbf_weights_generate: delays size " + std::to_string(delays_bytes) + " != expected " +
                    std::to_string(nd * sizeof(float2)));
    const long long total = (long long)nd * plan->F;
    FBF_REQUIRE(weights_bytes == (size_t)total * 2, FBF_ERR_SIZE,
                "fbf_weights_generate: weights size " + std::to_string(weights_bytes) + " != expected " +
                    std::to_string(total * 2));
    FBF_REQUIRE((reinterpret_cast<uintptr_t>(delays) & 7) == 0 && (reinterpret_cast<uintptr_t>(weights) & 1) == 0,
                FBF_ERR_ARG, "fbf_weights_generate: misaligned buffer");
    FBF_REQUIRE(nd < (1ull << 31), FBF_ERR_UNSUPPORTED, "fbf_weights_generate: too many (beam, antenna) pairs");
    const unsigned gy = (unsigned)((plan->F + kWChans - 1) / kWChans);
    FBF_REQUIRE(gy < 65536u, FBF_ERR_UNSUPPORTED, "fbf_weights_generate: too many channels");
    *out = WeightsLaunch{static_cast<const float2*>(delays), plan->d_freqs, epoch, plan->F, (int)nd,
                         static_cast<char2*>(weights), (unsigned)((nd + 255) / 256), gy};
    return FBF_OK;
}

extern "C" int fbf_weights_generate(fbf_weights_plan* plan, const void* delays, size_t delays_bytes, double epoch,
                                    void* weights, size_t weights_bytes, float* phase, fbf_stream_t stream)
{
    WeightsLaunch w{};
    const int st = weights_launch_params(plan, delays, delays_bytes, epoch, weights, weights_bytes, &w);
    if (st) return st;
    weights_k<<<dim3(w.gx, w.gy), 256, 0, static_cast<hipStream_t>(stream)>>>(w, phase);
    FBF_LAUNCH_CHECK("weights_k");
    return FBF_OK;
}

extern "C" int fbf_weights_generate_steps(fbf_weights_plan* plan, const void* delays, size_t delays_bytes,
                                          double tstart, double tstep, int ntsteps, void* weights,
                                          size_t weights_bytes, fbf_stream_t stream)
{
    FBF_REQUIRE(ntsteps > 0 && ntsteps <= 65535, FBF_ERR_ARG, "fbf_weights_generate_steps: ntsteps must be 1..65535");
    FBF_REQUIRE(weights_bytes % (size_t)ntsteps == 0, FBF_ERR_SIZE,
                "fbf_weights_generate_steps: weights size is not ntsteps whole steps");
    WeightsLaunch w{};
    const int st =
        weights_launch_params(plan, delays, delays_bytes, tstart, weights, weights_bytes / (size_t)ntsteps, &w);
    if (st) return st;
    weights_steps_k<<<dim3(w.gx, w.gy, (unsigned)ntsteps), 256, 0, static_cast<hipStream_t>(stream)>>>(w, tstart,
                                                                                                      tstep);
    FBF_LAUNCH_CHECK("weights_steps_k");
    return FBF_OK;
}
