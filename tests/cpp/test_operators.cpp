// C++ operator-layer tests, one per reference gtest on this path:
//   SplitTransposeTester.cycling_prime_test        (SplitTransposeTester.cu:96-113)
//   CoherentBeamformerTester.representative_noise_test (CoherentBeamformerTester.cu:148-185)
//   IncoherentBeamformerTester.ib_representative_noise_test (IncoherentBeamformerTester.cu:124-145)
//   WeightsManagerTester.test_zero_value / test_real_value (WeightsManagerTester.cu:95-119)
// plus DoubleDeviceBuffer / DeviceVector semantics.  Each compares the HIP
// path, called through the C++ classes, with the CPU oracle (oracle/).
// Exit status 0 = all passed.  Usage: fbfuse_cpp_tests [test-name-substring]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "fbf_oracle.h"
#include "psrdada_cpp/double_device_buffer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/CoherentBeamformer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/IncoherentBeamformer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"
#include "psrdada_cpp/meerkat/fbfuse/SplitTranspose.hpp"
#include "psrdada_cpp/meerkat/fbfuse/WeightsManager.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

using namespace psrdada_cpp;
using namespace psrdada_cpp::meerkat::fbfuse;

namespace {

int g_failures = 0;

#define EXPECT(cond)                                                                   \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::printf("    EXPECT failed at %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failures;                                                              \
            return;                                                                    \
        }                                                                              \
    } while (0)

std::vector<char2> gaussian_char2(uint64_t seed, double sigma, std::size_t n, bool clip)
{
    std::vector<char2> v(n);
    FBF_ERROR_CHECK(fbf_synth_gaussian_int8(seed, sigma, clip ? 1 : 0, reinterpret_cast<int8_t*>(v.data()), 2 * n));
    return v;
}

std::vector<char2> unit_weights(uint64_t seed, std::size_t n)
{
    std::vector<float> theta(n);
    FBF_ERROR_CHECK(fbf_synth_uniform_f32(seed, 0.0f, 6.2831853f, theta.data(), n));
    std::vector<char2> w(n);
    for (std::size_t i = 0; i < n; ++i) {
        w[i].x = (int8_t)std::lround(127.0 * std::cos((double)theta[i]));
        w[i].y = (int8_t)std::lround(127.0 * std::sin((double)theta[i]));
    }
    return w;
}

template <typename T>
const int8_t* bytes_of(std::vector<T> const& v)
{
    return reinterpret_cast<const int8_t*>(v.data());
}

// ---------------------------------------------------------------------------
void split_transpose_cycling_prime()
{
    PipelineConfig config;   // reference geometry: 64 antennas, 44 used, 64 channels
    SplitTranspose split_transpose(config);
    StreamType stream = nullptr;
    FBF_ERROR_CHECK(fbf_stream_create(&stream));
    const std::size_t ntimestamps = 12;
    const std::size_t n = ntimestamps * config.total_nantennas() * config.nchans() * config.nsamples_per_heap() *
                          config.npol();
    std::vector<char2> host(n);
    for (std::size_t ii = 0; ii < n; ++ii) {
        host[ii].x = (int8_t)(ii % 113);
        host[ii].y = (int8_t)(ii % 107);
    }
    SplitTranspose::VoltageType input(host), output;
    split_transpose.transpose(input, output, stream);
    std::vector<char2> got = output.to_host();
    std::vector<char2> ref(n / config.total_nantennas() * config.cb_nantennas());
    oracle_split_transpose(bytes_of(host), reinterpret_cast<int8_t*>(ref.data()), (int)ntimestamps,
                           (int)config.total_nantennas(), (int)config.cb_nantennas(),
                           (int)config.cb_antenna_offset(), (int)config.nchans(), (int)config.nsamples_per_heap(),
                           (int)config.npol());
    EXPECT(got.size() == ref.size());
    EXPECT(std::memcmp(got.data(), ref.data(), 2 * got.size()) == 0);
    fbf_stream_destroy(stream);
}

// ---------------------------------------------------------------------------
void coherent_beamformer_representative_noise()
{
    PipelineConfig config;   // 44 antennas, 1024 beams, tscrunch 16, fscrunch 4, 64 channels
    const float input_level = 32.0f;
    config.input_level(input_level);
    config.output_level(32.0f);
    CoherentBeamformer coherent_beamformer(config);
    StreamType stream = nullptr;
    FBF_ERROR_CHECK(fbf_stream_create(&stream));
    // CoherentBeamformerTester.cu:159-160: enough heap groups for one output heap.
    std::size_t ntimestamps = std::max<std::size_t>(
        1, 8192 / (config.nchans() / config.cb_fscrunch()) / (config.nsamples_per_heap() / config.cb_tscrunch()));
    const std::size_t nsamples = config.nsamples_per_heap() * ntimestamps;
    const std::size_t input_size = ntimestamps * config.cb_nantennas() * config.nchans() *
                                   config.nsamples_per_heap() * config.npol();
    const std::size_t weights_size = config.cb_nantennas() * config.nchans() * config.cb_nbeams();
    std::vector<char2> v = gaussian_char2(1001, input_level, input_size, false);
    std::vector<char2> w = unit_weights(1002, weights_size);
    CoherentBeamformer::VoltageVectorType d_v(v);
    CoherentBeamformer::WeightsVectorType d_w(w);
    CoherentBeamformer::PowerVectorType d_out;
    coherent_beamformer.beamform(d_v, d_w, d_out, stream);
    std::vector<int8_t> got = d_out.to_host();

    const int F = (int)config.nchans(), ts = (int)config.cb_tscrunch(), fs = (int)config.cb_fscrunch();
    const int A = (int)config.cb_nantennas(), B = (int)config.cb_nbeams();
    std::vector<double> power((std::size_t)(F / fs) * (nsamples / ts) * B);
    oracle_cb_power(bytes_of(v), bytes_of(w), power.data(), F, (long)nsamples, 2, A, B, ts, fs, 16);
    std::vector<float> off = config.cb_power_offsets(), sc = config.cb_power_scalings();
    std::vector<int8_t> ref(got.size());
    oracle_cb_tbtf(power.data(), ref.data(), off.data(), sc.data(), B, F / fs, (long)(nsamples / ts),
                   (int)config.cb_nsamples_per_heap(), config.quantisation_mode());
    EXPECT(got.size() == ref.size());
    std::size_t diff = 0;
    for (std::size_t i = 0; i < got.size(); ++i) diff += got[i] != ref[i];
    std::printf("    %zu outputs, %zu differ from the oracle\n", got.size(), diff);
    EXPECT(diff == 0);
    fbf_stream_destroy(stream);
}

// ---------------------------------------------------------------------------
void incoherent_beamformer_representative_noise()
{
    PipelineConfig config;   // 64 antennas, 64 channels, tscrunch 16, fscrunch 2
    const float input_level = 32.0f;
    config.input_level(input_level);
    config.output_level(32.0f);
    IncoherentBeamformer incoherent_beamformer(config);
    StreamType stream = nullptr;
    FBF_ERROR_CHECK(fbf_stream_create(&stream));
    const std::size_t ntimestamps = 32;
    const std::size_t input_size = ntimestamps * config.ib_nantennas() * config.nchans() *
                                   config.nsamples_per_heap() * config.npol();
    std::vector<char2> v = gaussian_char2(2001, input_level, input_size, false);
    IncoherentBeamformer::VoltageVectorType d_v(v);
    IncoherentBeamformer::PowerVectorType d_out;
    incoherent_beamformer.beamform(d_v, d_out, stream);
    std::vector<int8_t> got = d_out.to_host();
    const int F = (int)config.nchans(), ts = (int)config.ib_tscrunch(), fs = (int)config.ib_fscrunch();
    const long ntime = (long)(ntimestamps * config.nsamples_per_heap() / ts);
    std::vector<double> power((std::size_t)ntime * (F / fs));
    oracle_ib_power(bytes_of(v), power.data(), F, ts, fs, (int)ntimestamps, (int)config.ib_nantennas(), 2,
                    (int)config.nsamples_per_heap(), 16);
    std::vector<float> off = config.ib_power_offsets(), sc = config.ib_power_scalings();
    std::vector<int8_t> ref(got.size());
    oracle_ib_tf(power.data(), ref.data(), off.data(), sc.data(), ntime, F / fs, config.quantisation_mode());
    EXPECT(got.size() == ref.size());
    EXPECT(std::memcmp(got.data(), ref.data(), got.size()) == 0);
    // and within the reference's +-1 of its FP32 host model (IncoherentBeamformerTester.cu:120)
    std::vector<int8_t> ref32(got.size());
    oracle_ib_reference_f32(bytes_of(v), ref32.data(), F, ts, fs, (int)ntimestamps, (int)config.ib_nantennas(), 2,
                            (int)config.nsamples_per_heap(), config.ib_power_scaling(), config.ib_power_offset());
    for (std::size_t i = 0; i < got.size(); ++i) EXPECT(std::abs((int)got[i] - (int)ref32[i]) <= 1);
    fbf_stream_destroy(stream);
}

// ---------------------------------------------------------------------------
void weights_case(float rate, float offset, double epoch, bool exact_127)
{
    PipelineConfig config;
    config.centre_frequency(1.4e9f);
    config.bandwidth(56.0e6f);
    StreamType stream = nullptr;
    FBF_ERROR_CHECK(fbf_stream_create(&stream));
    WeightsManager weights_manager(config, stream);
    const std::size_t nd = config.cb_nbeams() * config.cb_nantennas();
    std::vector<float2> delays(nd, float2{rate, offset});
    WeightsManager::DelayVectorType d_delays(delays);
    auto const& weights = weights_manager.weights(d_delays, epoch);
    std::vector<char2> got = weights.to_host();
    std::vector<double> freqs = config.channel_frequencies_f64();
    std::vector<char2> ref(got.size());
    oracle_weights_f64(reinterpret_cast<const float*>(delays.data()), freqs.data(), epoch,
                       (int)config.cb_nantennas(), (int)config.cb_nbeams(), (int)config.nchans(),
                       reinterpret_cast<int8_t*>(ref.data()), nullptr);
    std::size_t off1 = 0;
    for (std::size_t i = 0; i < got.size(); ++i) {
        const int dx = std::abs(got[i].x - ref[i].x), dy = std::abs(got[i].y - ref[i].y);
        EXPECT(dx <= 1 && dy <= 1);
        off1 += (dx != 0) + (dy != 0);
        if (exact_127) EXPECT(got[i].x == 127 && got[i].y == 0);
    }
    std::printf("    %zu of %zu components differ by 1 LSB from the FP64 oracle\n", off1, 2 * got.size());
    EXPECT(off1 * 1000 < 2 * got.size());
    fbf_stream_destroy(stream);
}

void weights_zero_value() { weights_case(0.0f, 0.0f, 0.0, true); }
void weights_real_value() { weights_case(1e-11f, 1e-10f, 10.0, false); }

// ---------------------------------------------------------------------------
void double_device_buffer_semantics()
{
    DoubleDeviceBuffer<int8_t> db;
    db.resize(1000, 7);
    EXPECT(db.size() == 1000);
    int8_t* a0 = db.a_ptr();
    int8_t* b0 = db.b_ptr();
    std::vector<int8_t> x(1000, 3);
    db.a().assign(x.data(), x.size());
    db.swap();
    EXPECT(db.a_ptr() == b0 && db.b_ptr() == a0);
    std::vector<int8_t> got_b = db.b().to_host();
    std::vector<int8_t> got_a = db.a().to_host();
    EXPECT(got_b[999] == 3 && got_a[0] == 7);
    db.a().resize(1500);   // keeps the first 1000 elements, zero-fills the rest
    std::vector<int8_t> grown = db.a().to_host();
    EXPECT(grown.size() == 1500 && grown[999] == 7 && grown[1499] == 0);
    char buf[16];
    RawBytes rb(buf, 16, 4);
    EXPECT(rb.total_bytes() == 16 && rb.used_bytes() == 4 && !rb.on_device());
    rb.used_bytes(16);
    EXPECT(rb.used_bytes() == 16 && rb.ptr() == buf);
}

void error_paths_throw()
{
    PipelineConfig config;
    CoherentBeamformer cb(config);
    CoherentBeamformer::VoltageVectorType v(1000);   // not a whole number of samples
    CoherentBeamformer::WeightsVectorType w(10);
    CoherentBeamformer::PowerVectorType out;
    bool threw = false;
    try {
        cb.beamform(v, w, out, nullptr);
    } catch (std::runtime_error const& e) {
        threw = true;
        std::printf("    threw as expected: %.80s\n", e.what());
    }
    EXPECT(threw);
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string filter = argc > 1 ? argv[1] : "";
    struct Case {
        const char* name;
        std::function<void()> fn;
    };
    std::vector<Case> cases = {
        {"SplitTransposeTester.cycling_prime_test", split_transpose_cycling_prime},
        {"CoherentBeamformerTester.representative_noise_test", coherent_beamformer_representative_noise},
        {"IncoherentBeamformerTester.ib_representative_noise_test", incoherent_beamformer_representative_noise},
        {"WeightsManagerTester.test_zero_value", weights_zero_value},
        {"WeightsManagerTester.test_real_value", weights_real_value},
        {"DoubleDeviceBuffer.semantics", double_device_buffer_semantics},
        {"CoherentBeamformer.error_paths_throw", error_paths_throw},
    };
    int ran = 0;
    for (auto& c : cases) {
        if (!filter.empty() && std::string(c.name).find(filter) == std::string::npos) continue;
        std::printf("[ RUN      ] %s\n", c.name);
        std::fflush(stdout);
        const int before = g_failures;
        try {
            c.fn();
        } catch (std::exception const& e) {
            std::printf("    exception: %s\n", e.what());
            ++g_failures;
        }
        std::printf("%s %s\n", g_failures == before ? "[       OK ]" : "[  FAILED  ]", c.name);
        std::fflush(stdout);
        ++ran;
    }
    std::printf("%d tests, %d failures\n", ran, g_failures);
    return g_failures == 0 && ran > 0 ? 0 : 1;
}
