// Coherent beamformer entry point that allocates nothing.
//
// fbf_cb_beamform stages the int8 beams ([Fo][To][B]) in a buffer the plan
// allocates on first use; callers that cannot allocate at call time (code
// under stream capture, tight serving loops) use this entry point instead.
// Its staging buffer is a file-scope device array of fixed size, one copy per
// GPU the library runs on, held until the process ends; the kernels receive
// its device address exactly as they receive the caller's workspace in
// fbf_cb_beamform_ws.  Every call on one GPU shares that GPU's copy, so calls
// that overlap on one GPU (different streams, or threads) race.  The staging
// buffer needs no clearing: the beamformer writes every byte the reorder
// reads.
#include "fbf_internal.hpp"

#include <mutex>

namespace {

constexpr size_t kScratchBytes = 32u << 20;
__device__ __align__(256) int8_t scratch[kScratchBytes];

std::mutex g_mutex;
void* g_ptr[64] = {nullptr};

// The current device's copy of `scratch`, looked up on the first call on
// each device with that device current.  Host code uses only this address:
// the array's own name is a host-side stand-in, not a device address.
int scratch_ptr(void** out)
{
    int dev = 0;
    FBF_HIP_TRY(hipGetDevice(&dev));
    FBF_REQUIRE(dev >= 0 && dev < 64, FBF_ERR_UNSUPPORTED, "fbf_cb_beamform_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[dev]) {
        void* ptr = nullptr;
        FBF_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        FBF_REQUIRE(ptr, FBF_ERR_HIP, "fbf_cb_beamform_noalloc: no device address for the staging array");
        g_ptr[dev] = ptr;
    }
    *out = g_ptr[dev];
    return FBF_OK;
}

}  // namespace

extern "C" int fbf_cb_beamform_noalloc(fbf_cb_plan* plan, const void* ftpa, size_t ftpa_bytes,
                                       const void* fba_weights, size_t weights_bytes, void* tbtf, size_t tbtf_bytes,
                                       fbf_stream_t stream)
{
    FBF_REQUIRE(plan, FBF_ERR_ARG, "fbf_cb_beamform_noalloc: null plan");
    size_t staged = 0;
    int st = fbf_cb_output_bytes(plan, ftpa_bytes, &staged);
    if (st) return st;
    FBF_REQUIRE(staged <= kScratchBytes, FBF_ERR_SIZE,
                "fbf_cb_beamform_noalloc: the staged output (" + std::to_string(staged) +
                    " bytes) exceeds the library's fixed staging array (" + std::to_string(kScratchBytes) +
                    " bytes); use fbf_cb_beamform or fbf_cb_beamform_ws");
    void* ws = nullptr;
    st = scratch_ptr(&ws);
    if (st) return st;
    return fbf_cb_beamform_ws(plan, ftpa, ftpa_bytes, fba_weights, weights_bytes, tbtf, tbtf_bytes, ws,
                              kScratchBytes, stream);
}
