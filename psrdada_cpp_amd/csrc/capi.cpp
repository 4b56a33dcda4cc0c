// C ABI: error state, device plumbing, deterministic synthetic data and the
// configuration formulas of PipelineConfig.
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fbf_internal.hpp"

namespace fbf {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

int hip_status(hipError_t e, const char* what)
{
    return set_error(FBF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace fbf

using namespace fbf;

extern "C" const char* fbf_last_error(void) { return g_last_error.c_str(); }

extern "C" int fbf_version(void) { return 10000; }

// ---------------------------------------------------------------------------
// Device plumbing
// ---------------------------------------------------------------------------
extern "C" int fbf_device_count(int* count)
{
    FBF_REQUIRE(count, FBF_ERR_ARG, "fbf_device_count: null pointer");
    FBF_HIP_TRY(hipGetDeviceCount(count));
    return FBF_OK;
}

extern "C" int fbf_set_device(int device)
{
    FBF_HIP_TRY(hipSetDevice(device));
    return FBF_OK;
}

extern "C" int fbf_device_name(char* buf, size_t buflen)
{
    FBF_REQUIRE(buf && buflen > 0, FBF_ERR_ARG, "fbf_device_name: null buffer");
    int dev = 0;
    FBF_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FBF_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    // The marketing name can be empty (it comes from the driver's product
    // table); the architecture name and CU count are always there.
    std::string s = std::string(prop.gcnArchName) + ", " + std::to_string(prop.multiProcessorCount) + " CUs, " +
                    std::to_string(prop.totalGlobalMem >> 30) + " GiB";
    if (prop.name[0]) s = std::string(prop.name) + " (" + s + ")";
    std::strncpy(buf, s.c_str(), buflen - 1);
    buf[buflen - 1] = 0;
    return FBF_OK;
}

extern "C" int fbf_malloc(void** ptr, size_t bytes)
{
    FBF_REQUIRE(ptr, FBF_ERR_ARG, "fbf_malloc: null pointer");
    *ptr = nullptr;
    FBF_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return FBF_OK;
}

extern "C" int fbf_free(void* ptr)
{
    if (!ptr) return FBF_OK;
    FBF_HIP_TRY(hipFree(ptr));
    return FBF_OK;
}

extern "C" int fbf_host_malloc(void** ptr, size_t bytes)
{
    FBF_REQUIRE(ptr, FBF_ERR_ARG, "fbf_host_malloc: null pointer");
    *ptr = nullptr;
    FBF_HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return FBF_OK;
}

extern "C" int fbf_host_free(void* ptr)
{
    if (!ptr) return FBF_OK;
    FBF_HIP_TRY(hipHostFree(ptr));
    return FBF_OK;
}

extern "C" int fbf_memcpy_h2d(void* dst, const void* src, size_t bytes, fbf_stream_t stream)
{
    if (!bytes) return FBF_OK;
    FBF_REQUIRE(dst && src, FBF_ERR_ARG, "fbf_memcpy_h2d: null pointer");
    FBF_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_memcpy_d2h(void* dst, const void* src, size_t bytes, fbf_stream_t stream)
{
    if (!bytes) return FBF_OK;
    FBF_REQUIRE(dst && src, FBF_ERR_ARG, "fbf_memcpy_d2h: null pointer");
    FBF_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_memcpy_d2d(void* dst, const void* src, size_t bytes, fbf_stream_t stream)
{
    if (!bytes) return FBF_OK;
    FBF_REQUIRE(dst && src, FBF_ERR_ARG, "fbf_memcpy_d2d: null pointer");
    FBF_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_memset(void* dst, int value, size_t bytes, fbf_stream_t stream)
{
    if (!bytes) return FBF_OK;
    FBF_REQUIRE(dst, FBF_ERR_ARG, "fbf_memset: null pointer");
    FBF_HIP_TRY(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_stream_create(fbf_stream_t* stream)
{
    FBF_REQUIRE(stream, FBF_ERR_ARG, "fbf_stream_create: null pointer");
    hipStream_t s = nullptr;
    FBF_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return FBF_OK;
}

extern "C" int fbf_stream_destroy(fbf_stream_t stream)
{
    if (!stream) return FBF_OK;
    FBF_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_stream_synchronize(fbf_stream_t stream)
{
    FBF_HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_device_synchronize(void)
{
    FBF_HIP_TRY(hipDeviceSynchronize());
    return FBF_OK;
}

extern "C" int fbf_event_create(fbf_event_t* event)
{
    FBF_REQUIRE(event, FBF_ERR_ARG, "fbf_event_create: null pointer");
    hipEvent_t e = nullptr;
    FBF_HIP_TRY(hipEventCreate(&e));
    *event = e;
    return FBF_OK;
}

extern "C" int fbf_event_destroy(fbf_event_t event)
{
    if (!event) return FBF_OK;
    FBF_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return FBF_OK;
}

extern "C" int fbf_event_record(fbf_event_t event, fbf_stream_t stream)
{
    FBF_REQUIRE(event, FBF_ERR_ARG, "fbf_event_record: null event");
    FBF_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return FBF_OK;
}

extern "C" int fbf_stream_wait_event(fbf_stream_t stream, fbf_event_t event)
{
    FBF_REQUIRE(event, FBF_ERR_ARG, "fbf_stream_wait_event: null event");
    FBF_HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(stream), static_cast<hipEvent_t>(event), 0));
    return FBF_OK;
}

extern "C" int fbf_event_synchronize(fbf_event_t event)
{
    FBF_REQUIRE(event, FBF_ERR_ARG, "fbf_event_synchronize: null event");
    FBF_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return FBF_OK;
}

extern "C" int fbf_event_elapsed_ms(fbf_event_t start, fbf_event_t stop, float* ms)
{
    FBF_REQUIRE(start && stop && ms, FBF_ERR_ARG, "fbf_event_elapsed_ms: null argument");
    FBF_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return FBF_OK;
}

// ---------------------------------------------------------------------------
// Deterministic synthetic data.  The k-th uniform of a stream with seed s is
// splitmix64(s + (k + 1) * 0x9E3779B97F4A7C15) >> 11 scaled by 2^-53, i.e.
// the splitmix64 sequence evaluated counter-style so that it can be filled in
// parallel.  Gaussians use Box-Muller on uniforms 2i and 2i+1.
// ---------------------------------------------------------------------------
namespace {

inline uint64_t splitmix_at(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double unit_at(uint64_t seed, uint64_t k)
{
    return (double)(splitmix_at(seed, k) >> 11) * (1.0 / 9007199254740992.0);
}

template <class F>
void parallel_for(size_t n, F&& body)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < (1u << 16)) nt = 1;
    std::vector<std::thread> pool;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([lo, hi, &body] { body(lo, hi); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" int fbf_synth_gaussian_int8(uint64_t seed, double sigma, int clip, int8_t* out, size_t n)
{
    FBF_REQUIRE(out || n == 0, FBF_ERR_ARG, "fbf_synth_gaussian_int8: null output");
    const double two_pi = 6.283185307179586476925286766559;
    const size_t npairs = (n + 1) / 2;
    parallel_for(npairs, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const double u1 = 1.0 - unit_at(seed, 2 * i);  // (0, 1]
            const double u2 = unit_at(seed, 2 * i + 1);
            const double r = std::sqrt(-2.0 * std::log(u1));
            const double z[2] = {r * std::cos(two_pi * u2), r * std::sin(two_pi * u2)};
            for (int h = 0; h < 2; ++h) {
                const size_t idx = 2 * i + h;
                if (idx >= n) break;
                long v = std::lround(sigma * z[h]);
                if (clip) {
                    if (v > 127) v = 127;
                    if (v < -127) v = -127;
                }
                out[idx] = (int8_t)(uint8_t)(v & 0xFF);
            }
        }
    });
    return FBF_OK;
}

extern "C" int fbf_synth_uniform_f32(uint64_t seed, float lo, float hi, float* out, size_t n)
{
    FBF_REQUIRE(out || n == 0, FBF_ERR_ARG, "fbf_synth_uniform_f32: null output");
    parallel_for(n, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i)
            out[i] = (float)((double)lo + ((double)hi - (double)lo) * unit_at(seed, i));
    });
    return FBF_OK;
}

// ---------------------------------------------------------------------------
// PipelineConfig formulas
// ---------------------------------------------------------------------------
extern "C" int fbf_power_levels(float input_level, float output_level, int cb_nantennas, int cb_tscrunch,
                                int cb_fscrunch, int ib_nantennas, int ib_tscrunch, int ib_fscrunch, int npol,
                                float* cb_offset, float* cb_scale, float* ib_offset, float* ib_scale)
{
#pragma clang fp contract(off)
    FBF_REQUIRE(cb_offset && cb_scale && ib_offset && ib_scale, FBF_ERR_ARG, "fbf_power_levels: null output");
    FBF_REQUIRE(output_level != 0.0f, FBF_ERR_ARG, "fbf_power_levels: output level must be non-zero");
    // PipelineConfig.cpp:112-118 (coherent beam)
    const float weights_amp = 127.0f;
    const float amp = weights_amp * input_level * std::sqrt(static_cast<float>(cb_nantennas));
    const float cb_s = static_cast<float>(std::pow(static_cast<double>(amp), 2));
    const float cb_dof = static_cast<float>(2 * (size_t)cb_tscrunch * (size_t)cb_fscrunch * (size_t)npol);
    *cb_offset = cb_s * cb_dof;
    *cb_scale = cb_s * std::sqrt(2 * cb_dof) / output_level;
    // PipelineConfig.cpp:120-124 (incoherent beam)
    const float ib_s = static_cast<float>(std::pow(static_cast<double>(input_level), 2));
    const float ib_dof = static_cast<float>(2 * (size_t)ib_tscrunch * (size_t)ib_fscrunch * (size_t)ib_nantennas *
                                            (size_t)npol);
    *ib_offset = ib_s * ib_dof;
    *ib_scale = ib_s * std::sqrt(2 * ib_dof) / output_level;
    return FBF_OK;
}

extern "C" int fbf_channel_frequencies(float cfreq, float bw, int nchans, float* freqs)
{
#pragma clang fp contract(off)
    FBF_REQUIRE(freqs && nchans > 0, FBF_ERR_ARG, "fbf_channel_frequencies: bad argument");
    // PipelineConfig.cpp:184-193
    const float chbw = bw / static_cast<float>(nchans);
    const float fbottom = cfreq - bw / 2.0f;
    for (int c = 0; c < nchans; ++c) freqs[c] = fbottom + chbw / 2.0f + (chbw * static_cast<float>(c));
    return FBF_OK;
}

extern "C" int fbf_channel_frequencies_f64(double cfreq, double bw, int nchans, double* freqs)
{
    FBF_REQUIRE(freqs && nchans > 0, FBF_ERR_ARG, "fbf_channel_frequencies_f64: bad argument");
    const double chbw = bw / nchans;
    const double fbottom = cfreq - bw / 2.0;
    for (int c = 0; c < nchans; ++c) freqs[c] = fbottom + chbw / 2.0 + chbw * c;
    return FBF_OK;
}
