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
#include <thread>
#include <utility>

#include <sys/wait.h>
#include <unistd.h>

#include "psrdada_cpp/dada_input_stream.hpp"
#include "psrdada_cpp/dada_write_client.hpp"
#include "psrdada_cpp/double_device_buffer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/DelayManager.hpp"
#include "psrdada_cpp/meerkat/fbfuse/Header.hpp"
#include "psrdada_cpp/meerkat/fbfuse/Pipeline.hpp"
#include "psrdada_cpp/meerkat/fbfuse/CoherentBeamformer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/IncoherentBeamformer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"
#include "psrdada_cpp/meerkat/fbfuse/SplitTranspose.hpp"
#include "psrdada_cpp/meerkat/fbfuse/WeightsManager.hpp"
#include "psrdada_cpp/meerkat/tools/feng_to_dada.hpp"
#include "psrdada_cpp/meerkat/tuse/transpose_to_dada.hpp"
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

// ---------------------------------------------------------------------------
void header_get_set()
{
    std::vector<char> buf(4096, 0);
    RawBytes rb(buf.data(), buf.size(), buf.size());
    Header h(rb);
    h.set<long double>("SAMPLE_CLOCK", 856000000.0L);
    h.set<long double>("SYNC_TIME", 1234.5L);
    h.set<std::size_t>("SAMPLE_CLOCK_START", 42);
    EXPECT(h.get<long double>("SAMPLE_CLOCK") == 856000000.0L);
    EXPECT(h.get<long double>("SYNC_TIME") == 1234.5L);
    EXPECT(h.get<std::size_t>("SAMPLE_CLOCK_START") == 42);
    h.set<std::size_t>("SAMPLE_CLOCK_START", 7);   // replace, not append
    EXPECT(h.get<std::size_t>("SAMPLE_CLOCK_START") == 7);
    EXPECT(std::string(buf.data()).find("SAMPLE_CLOCK_START 42") == std::string::npos);
    bool threw = false;
    try {
        h.get<long double>("NOT_A_KEY");
    } catch (std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw);
    h.set_string("UNSET_KEY", "unset");
    threw = false;
    try {
        h.get<long double>("UNSET_KEY");
    } catch (std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw);
    h.purge();
    EXPECT(buf[0] == 0);
}

void dada_ring_write_read_eod()
{
    DadaDB db(4, 64, 2, 256);
    db.create();
    MultiLog log("test");
    DadaWriteClient writer(db.key(), log);
    std::thread producer([&] {
        auto& hb = writer.header_stream().next();
        std::strcpy(hb.ptr(), "HDR 1\n");
        hb.used_bytes(7);
        writer.header_stream().release();
        for (int i = 0; i < 10; ++i) {
            auto& b = writer.data_stream().next();
            std::memset(b.ptr(), i, 64);
            b.used_bytes(64 - i);
            writer.data_stream().release(i == 9);
        }
    });
    DadaReadClient reader(db.key(), log);
    auto& h = reader.header_stream().next();
    EXPECT(std::string(h.ptr()) == "HDR 1\n");
    reader.header_stream().release();
    int n = 0;
    while (true) {
        auto& b = reader.data_stream().next();
        EXPECT(b.used_bytes() == (std::size_t)(64 - n) && b.ptr()[0] == (char)n);
        reader.data_stream().release();
        ++n;
        if (reader.data_stream().at_end()) break;
    }
    producer.join();
    EXPECT(n == 10);
    bool threw = false;
    try {
        DadaWriteClient second(db.key(), log);   // one writer per ring
    } catch (std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw);
}

// A writer in another process fills two observations of a shared ring
// (SysV segments key and key + 1, as dada_db.cpp:95-110 creates them); this
// process reads them through DadaInputStream, which attaches a new read
// client for each observation (detail/dada_input_stream.cpp:29-50).  The
// writer is forked before anything touches a device.
struct ObservationRecorder {
    std::vector<std::string> headers;
    std::vector<std::pair<int, int>> blocks;   // (observation, block) from the payload
    std::vector<std::size_t> used;
    bool payload_ok = true;
    std::size_t stop_after = 0;
    void init(RawBytes& h) { headers.emplace_back(h.ptr(), strnlen(h.ptr(), h.used_bytes())); }
    bool operator()(RawBytes& b)
    {
        const int o = b.ptr()[0], i = b.ptr()[1];
        for (std::size_t k = 2; k < b.used_bytes(); ++k)
            if (b.ptr()[k] != (char)(o * 31 + i * 7 + k)) payload_ok = false;
        blocks.emplace_back(o, i);
        used.push_back(b.used_bytes());
        return blocks.size() == stop_after;
    }
};

void dada_ring_cross_process()
{
    DadaDB db(3, 4096, 2, 4096);   // shared placement (the default)
    db.create();
    const key_t key = db.key();
    const int nobs = 2, nblk = 5;
    std::fflush(stdout);
    const pid_t pid = fork();
    if (pid == 0) {
        int rc = 0;
        try {
            MultiLog log("writer");
            DadaWriteClient writer(key, log);
            for (int o = 0; o < nobs; ++o) {
                if (o) writer.reset();   // the next observation
                auto& h = writer.header_stream().next();
                std::snprintf(h.ptr(), h.total_bytes(), "OBS %d\nSOURCE J0437-4715\n", o);
                h.used_bytes(std::strlen(h.ptr()) + 1);
                writer.header_stream().release();
                for (int i = 0; i < nblk; ++i) {
                    auto& b = writer.data_stream().next();
                    const std::size_t used = b.total_bytes() - 64 * i;
                    b.ptr()[0] = (char)o;
                    b.ptr()[1] = (char)i;
                    for (std::size_t k = 2; k < used; ++k) b.ptr()[k] = (char)(o * 31 + i * 7 + k);
                    b.used_bytes(used);
                    writer.data_stream().release(i == nblk - 1);
                }
            }
        } catch (std::exception const& e) {
            std::fprintf(stderr, "writer: %s\n", e.what());
            rc = 1;
        }
        _exit(rc);
    }
    EXPECT(pid > 0);
    MultiLog log("reader");
    ObservationRecorder rec;
    rec.stop_after = (std::size_t)(nobs * nblk);
    DadaInputStream<ObservationRecorder> stream(key, log, rec);
    stream.start();
    int status = 0;
    EXPECT(waitpid(pid, &status, 0) == pid);
    EXPECT(WIFEXITED(status) && WEXITSTATUS(status) == 0);
    EXPECT(rec.headers.size() == (std::size_t)nobs);
    EXPECT(rec.headers[0].rfind("OBS 0\n", 0) == 0 && rec.headers[1].rfind("OBS 1\n", 0) == 0);
    EXPECT(rec.blocks.size() == (std::size_t)(nobs * nblk) && rec.payload_ok);
    for (int o = 0; o < nobs; ++o)
        for (int i = 0; i < nblk; ++i) {
            const std::size_t k = (std::size_t)(o * nblk + i);
            EXPECT(rec.blocks[k] == std::make_pair(o, i) && rec.used[k] == 4096 - 64 * (std::size_t)i);
        }
}

// One writer per ring across processes: a second writer in this process is
// refused while the writer process holds the lock, and takes over a lock
// left by a writer process that has exited without releasing it.
void dada_ring_writer_lock_across_processes()
{
    DadaDB db(2, 256, 2, 256);
    db.create();
    const key_t key = db.key();
    int to_child[2], to_parent[2];
    EXPECT(pipe(to_child) == 0 && pipe(to_parent) == 0);
    std::fflush(stdout);
    const pid_t pid = fork();
    if (pid == 0) {
        MultiLog log("writer");
        DadaWriteClient* writer = nullptr;
        char c = 'x';
        try {
            writer = new DadaWriteClient(key, log);   // never released: the process exits holding it
            c = 'L';
        } catch (...) {
        }
        (void)!write(to_parent[1], &c, 1);
        (void)!read(to_child[0], &c, 1);   // wait until the parent has tried
        _exit(writer ? 0 : 1);
    }
    EXPECT(pid > 0);
    char c = 0;
    EXPECT(read(to_parent[0], &c, 1) == 1 && c == 'L');
    MultiLog log("second");
    bool refused = false;
    try {
        DadaWriteClient second(key, log);
    } catch (std::runtime_error const&) {
        refused = true;
    }
    EXPECT(refused);
    (void)!write(to_child[1], "g", 1);
    int status = 0;
    EXPECT(waitpid(pid, &status, 0) == pid && WIFEXITED(status) && WEXITSTATUS(status) == 0);
    DadaWriteClient taker(key, log);   // the exited writer's lock is taken over
    auto& b = taker.data_stream().next();
    b.used_bytes(1);
    taker.data_stream().release();
    for (int fd : {to_child[0], to_child[1], to_parent[0], to_parent[1]}) close(fd);
}

// DelayManagerTester.cu:47-67
void delay_manager_updates()
{
    PipelineConfig config;
    config.delay_buffer_shm("fbfuse_test_delays_shm");
    config.delay_buffer_mutex("fbfuse_test_delays_mutex");
    config.delay_buffer_sem("fbfuse_test_delays_count");
    DelayEngineSimulator simulator(config);
    StreamType stream = nullptr;
    FBF_ERROR_CHECK(fbf_stream_create(&stream));
    DelayManager delay_manager(config, stream);
    const std::size_t n = config.cb_nbeams() * config.cb_nantennas();
    std::memset(simulator.delays(), 0x01, n * sizeof(float2));
    simulator.header()->epoch = 10.0;
    simulator.header()->duration = 5.0;
    simulator.update_delays();
    auto const& d = delay_manager.delays();
    std::vector<float2> got = d.to_host();
    EXPECT(std::memcmp(got.data(), simulator.delays(), n * sizeof(float2)) == 0);
    EXPECT(delay_manager.epoch() == 10.0 && delay_manager.duration() == 5.0 && delay_manager.update_count() == 1);
    delay_manager.delays();   // no new update: no copy
    EXPECT(delay_manager.update_count() == 1);
    std::memset(simulator.delays(), 0x02, n * sizeof(float2));
    simulator.update_delays();
    got = delay_manager.delays().to_host();
    EXPECT(std::memcmp(got.data(), simulator.delays(), n * sizeof(float2)) == 0 && delay_manager.update_count() == 2);
    fbf_stream_destroy(stream);
}

void delay_manager_bad_keys()
{
    PipelineConfig config;
    config.delay_buffer_shm("fbfuse_test_no_such_shm");
    bool threw = false;
    try {
        DelayManager delay_manager(config, nullptr);
    } catch (std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw);
}

// PipelineTester.cu:46-125: rings, null-sink consumers on their own threads,
// a synthetic header, blocks pushed through, stop mid-stream.
void pipeline_simple_run()
{
    PipelineConfig config;   // reference geometry
    config.delay_buffer_shm("fbfuse_test_pipe_shm");
    config.delay_buffer_mutex("fbfuse_test_pipe_mutex");
    config.delay_buffer_sem("fbfuse_test_pipe_count");
    config.centre_frequency(1.4e9f);
    config.bandwidth(56e6f);
    DelayEngineSimulator simulator(config);
    const int ntimestamps_per_block = 64;
    const std::size_t taftp_block_bytes = (std::size_t)ntimestamps_per_block * config.total_nantennas() *
                                          config.nchans() * config.nsamples_per_heap() * config.npol() * sizeof(char2);
    const std::size_t cb_block = config.cb_nbeams() * (config.nsamples_per_heap() * ntimestamps_per_block /
                                                        config.cb_tscrunch()) *
                                 (config.nchans() / config.cb_fscrunch());
    DadaDB cb_buffer(8, cb_block, 4, 4096);
    cb_buffer.create();
    config.cb_dada_key(cb_buffer.key());
    const std::size_t ib_block = (config.nsamples_per_heap() * ntimestamps_per_block / config.ib_tscrunch()) *
                                 (config.nchans() / config.ib_fscrunch());
    DadaDB ib_buffer(8, ib_block, 4, 4096);
    ib_buffer.create();
    config.ib_dada_key(ib_buffer.key());
    MultiLog log("PipelineTester");
    DadaWriteClient cb_write_client(config.cb_dada_key(), log);
    DadaWriteClient ib_write_client(config.ib_dada_key(), log);
    std::size_t seen_cb = 0;
    {
        Pipeline pipeline(config, cb_write_client, ib_write_client, taftp_block_bytes);
        NullSink null_sink;
        DadaInputStream<NullSink> cb_consumer(config.cb_dada_key(), log, null_sink);
        DadaInputStream<NullSink> ib_consumer(config.ib_dada_key(), log, null_sink);
        std::thread cb_thread([&] { cb_consumer.start(); });
        std::thread ib_thread([&] { ib_consumer.start(); });
        std::vector<char> header_buffer(4096, 0);
        RawBytes header_rb(header_buffer.data(), 4096, 4096);
        Header header(header_rb);
        header.set<long double>("SAMPLE_CLOCK", 856000000.0L);
        header.set<long double>("SYNC_TIME", 0.0L);
        header.set<std::size_t>("SAMPLE_CLOCK_START", 0);
        void* input = nullptr;
        FBF_ERROR_CHECK(fbf_host_malloc(&input, taftp_block_bytes));
        std::memset(input, 0, taftp_block_bytes);
        RawBytes input_rb(static_cast<char*>(input), taftp_block_bytes, taftp_block_bytes);
        pipeline.init(header_rb);
        for (int ii = 0; ii < 20; ++ii) {
            if (ii == 10) simulator.update_delays();   // delay model refresh mid-stream
            pipeline(input_rb);
        }
        cb_consumer.stop();
        ib_consumer.stop();
        pipeline(input_rb);   // one more block wakes the consumers
        pipeline(input_rb);
        cb_thread.join();
        ib_thread.join();
        seen_cb = pipeline.call_count();
        // timestamp of the block processed by call n: (n - 2) * tick / clock
        const long double tick = 2.0L * config.total_nchans() * config.nsamples_per_heap() * ntimestamps_per_block;
        EXPECT(std::fabs((double)(pipeline.unix_timestamp() - (seen_cb - 2) * tick / 856000000.0L)) < 1e-9);
        FBF_ERROR_CHECK(fbf_host_free(input));
    }
    EXPECT(seen_cb == 22);
}

// The beams leaving the pipeline are those of the operators for the same
// block (block n in, its beams out on call n + 2), with the delay model read
// from shared memory and weights at each block's unix time: coherent beams
// through SplitTranspose + CoherentBeamformer, the incoherent beam through
// IncoherentBeamformer.  At 16 antennas the pipeline split-transposes (the
// reference's order); at 64 it forms both beams in one pass over the TAFTP
// block, each block's reorder launch generating the next block's weights;
// for 64 antennas of a 128-antenna block it forms the coherent beams
// straight from the TAFTP block, the weights in the incoherent beam's launch.
// The delay model changes before the call that processes block 2, whose
// weights the one-pass route had already generated from the old model: the
// pipeline must generate them again.
struct CaptureSink {
    std::vector<std::vector<int8_t>> blocks;
    std::string header;
    void init(RawBytes& h) { header.assign(h.ptr(), strnlen(h.ptr(), h.total_bytes())); }
    bool operator()(RawBytes& b)
    {
        blocks.emplace_back(b.ptr(), b.ptr() + b.used_bytes());
        return false;
    }
};

void pipeline_outputs_match_operators(int Atot, int A, int a0, int B)
{
    PipelineConfig config;
    config.total_nantennas(Atot);
    config.nchans(16);
    config.cb_nantennas(A);
    config.cb_antenna_offset(a0);
    config.cb_nbeams(B);
    config.cb_fscrunch(1);
    config.cb_tscrunch(16);
    config.ib_fscrunch(1);
    config.cb_heap_size(2048);   // 128 output samples x 16 channels per heap
    config.input_level(16.0f);
    config.quantisation_mode(FBF_QUANT_ROUND);
    config.centre_frequency(1284e6f);
    config.bandwidth(856e6f);
    config.delay_buffer_shm("fbfuse_test_pv_shm");
    config.delay_buffer_mutex("fbfuse_test_pv_mutex");
    config.delay_buffer_sem("fbfuse_test_pv_count");
    DelayEngineSimulator simulator(config);
    const std::size_t nd = config.cb_nbeams() * config.cb_nantennas();
    std::vector<float> offs(nd);
    FBF_ERROR_CHECK(fbf_synth_uniform_f32(77, -1e-6f, 1e-6f, offs.data(), nd));
    for (std::size_t i = 0; i < nd; ++i) simulator.delays()[i] = float2{1e-9f, offs[i]};
    simulator.update_delays();
    std::vector<float2> delays_before(simulator.delays(), simulator.delays() + nd);
    std::vector<float> offs2(nd);
    FBF_ERROR_CHECK(fbf_synth_uniform_f32(78, -1e-6f, 1e-6f, offs2.data(), nd));

    const int nhg = 8;
    const std::size_t block_bytes = (std::size_t)nhg * Atot * 16 * 256 * 2 * 2;
    const std::size_t cb_bytes = (std::size_t)B * (nhg * 256 / 16) * 16;
    const std::size_t ib_bytes = (nhg * 256 / 16) * 16;
    DadaDB cb_buffer(16, cb_bytes, 4, 4096), ib_buffer(16, ib_bytes, 4, 4096);
    cb_buffer.create();
    ib_buffer.create();
    MultiLog log("PipelineValue");
    DadaWriteClient cbw(cb_buffer.key(), log), ibw(ib_buffer.key(), log);
    const int nblocks = 6;
    std::vector<std::vector<char2>> inputs;
    for (int i = 0; i < nblocks; ++i) inputs.push_back(gaussian_char2(5000 + i, 16.0, block_bytes / 2, true));
    CaptureSink cb_sink, ib_sink;
    long double ts[nblocks];
    {
        Pipeline pipeline(config, cbw, ibw, block_bytes);
        std::vector<char> hbuf(4096, 0);
        RawBytes hrb(hbuf.data(), 4096, 4096);
        Header header(hrb);
        header.set<long double>("SAMPLE_CLOCK", 1712000000.0L);
        header.set<long double>("SYNC_TIME", 1600000000.0L);
        header.set<std::size_t>("SAMPLE_CLOCK_START", 123456);
        pipeline.init(hrb);
        void* in = nullptr;
        FBF_ERROR_CHECK(fbf_host_malloc(&in, block_bytes));
        RawBytes in_rb(static_cast<char*>(in), block_bytes, block_bytes);
        for (int i = 0; i < nblocks; ++i) {
            std::memcpy(in, inputs[i].data(), block_bytes);
            if (i == 3) {   // call 3 processes block 2
                for (std::size_t j = 0; j < nd; ++j) simulator.delays()[j] = float2{-2e-9f, offs2[j]};
                simulator.update_delays();
            }
            pipeline(in_rb);
            if (i >= 1) ts[i - 1] = pipeline.unix_timestamp();
        }
        FBF_ERROR_CHECK(fbf_host_free(in));
    }   // destructor releases the last output block
    {
        DadaReadClient reader(cb_buffer.key(), log);
        auto& h = reader.header_stream().next();
        cb_sink.init(h);
        reader.header_stream().release();
        for (int i = 0; i < nblocks - 2; ++i) {
            cb_sink(reader.data_stream().next());
            reader.data_stream().release();
        }
    }
    {
        DadaReadClient reader(ib_buffer.key(), log);
        auto& h = reader.header_stream().next();
        ib_sink.init(h);
        reader.header_stream().release();
        for (int i = 0; i < nblocks - 2; ++i) {
            ib_sink(reader.data_stream().next());
            reader.data_stream().release();
        }
    }
    EXPECT(cb_sink.header.find("SAMPLE_CLOCK_START 123456") != std::string::npos);
    EXPECT(cb_sink.blocks.size() == (std::size_t)(nblocks - 2));
    // reference result per block through the operator classes
    StreamType stream = nullptr;
    FBF_ERROR_CHECK(fbf_stream_create(&stream));
    SplitTranspose st(config);
    CoherentBeamformer cb(config);
    IncoherentBeamformer ib(config);
    WeightsManager wm(config, stream);
    EXPECT(ib_sink.blocks.size() == (std::size_t)(nblocks - 2));
    std::vector<float2> dh(simulator.delays(), simulator.delays() + nd);
    WeightsManager::DelayVectorType dd_before(delays_before), dd_after(dh);
    for (int i = 0; i < nblocks - 2; ++i) {
        SplitTranspose::VoltageType taftp(inputs[i]), ftpa;
        st.transpose(taftp, ftpa, stream);
        auto const& w = wm.weights(i < 2 ? dd_before : dd_after, (double)ts[i]);
        CoherentBeamformer::PowerVectorType out;
        cb.beamform(ftpa, w, out, stream);
        std::vector<int8_t> expect = out.to_host();
        EXPECT(expect.size() == cb_sink.blocks[i].size());
        EXPECT(std::memcmp(expect.data(), cb_sink.blocks[i].data(), expect.size()) == 0);
        IncoherentBeamformer::PowerVectorType ib_out;
        ib.beamform(taftp, ib_out, stream);
        std::vector<int8_t> ib_expect = ib_out.to_host();
        EXPECT(ib_expect.size() == ib_sink.blocks[i].size());
        EXPECT(std::memcmp(ib_expect.data(), ib_sink.blocks[i].data(), ib_expect.size()) == 0);
    }
    // the epochs advance by one block of samples
    EXPECT(std::fabs((double)(ts[1] - ts[0]) - 2.0 * 4096 * nhg * 256 / 1712e6) < 1e-9);
    fbf_stream_destroy(stream);
}

// A DADA-handler stand-in that keeps what it was given.
struct BytesSink {
    std::vector<char> header;
    std::vector<char> data;
    int calls = 0;
    void init(RawBytes& block) { header.assign(block.ptr(), block.ptr() + block.used_bytes()); }
    bool operator()(RawBytes& block)
    {
        data.assign(block.ptr(), block.ptr() + block.used_bytes());
        ++calls;
        return false;
    }
};

// TransposeToDada (meerkat/tuse/transpose_to_dada.hpp) with the reference's
// handler-per-beam use (transpose_to_dada_cli.cpp:99-109): every beam's
// buffer equals the oracle's do_transpose for that beam; do_transpose itself
// on host and on device buffers; the reference's default geometry.
void tuse_transpose_to_dada()
{
    const int nbeams = 6, nchans = 64, nsamples = 32, nfreq = 8, ngroups = 3;
    std::vector<std::shared_ptr<BytesSink>> sinks;
    for (int i = 0; i < nbeams; ++i) sinks.push_back(std::make_shared<BytesSink>());
    meerkat::tuse::TransposeToDada<BytesSink> tr(nbeams, sinks);
    EXPECT(tr.nchans() == 128 && tr.nsamples() == 64 && tr.nfreq() == 32 && tr.ngroups() == 10 && tr.nbeams() == 6);
    tr.set_nchans(nchans);
    tr.set_nsamples(nsamples);
    tr.set_nfreq(nfreq);
    tr.set_ngroups(ngroups);
    const std::size_t per_beam = (std::size_t)nchans * nsamples * nfreq * ngroups;
    std::vector<char2> v = gaussian_char2(4242, 40.0, per_beam * nbeams / 2, false);
    char* blk = reinterpret_cast<char*>(v.data());
    std::vector<char> hdr(4096, 'h');
    RawBytes hrb(hdr.data(), hdr.size(), hdr.size());
    tr.init(hrb);
    RawBytes block(blk, per_beam * nbeams, per_beam * nbeams);
    EXPECT(tr(block) == false);
    std::vector<int8_t> expect(per_beam);
    for (int b = 0; b < nbeams; ++b) {
        oracle_tuse_transpose(reinterpret_cast<const int8_t*>(blk), expect.data(), nchans, nsamples, nfreq, nbeams,
                              ngroups, b, 1);
        EXPECT(sinks[b]->calls == 1 && sinks[b]->header.size() == 4096);
        EXPECT(sinks[b]->data.size() == per_beam);
        EXPECT(std::memcmp(sinks[b]->data.data(), expect.data(), per_beam) == 0);
    }
    // do_transpose for one beam, host buffers and device buffers
    std::vector<char> one(per_beam);
    RawBytes out_rb(one.data(), per_beam);
    meerkat::tuse::transpose::do_transpose(out_rb, block, nchans, nsamples, nfreq, 4, nbeams, ngroups);
    oracle_tuse_transpose(reinterpret_cast<const int8_t*>(blk), expect.data(), nchans, nsamples, nfreq, nbeams,
                          ngroups, 4, 1);
    EXPECT(std::memcmp(one.data(), expect.data(), per_beam) == 0);
    DeviceVector<char> din(std::vector<char>(blk, blk + per_beam * nbeams)), dout(per_beam);
    RawBytes din_rb(din.data(), din.bytes(), din.bytes(), true), dout_rb(dout.data(), dout.bytes(), 0, true);
    meerkat::tuse::transpose::do_transpose(dout_rb, din_rb, nchans, nsamples, nfreq, 4, nbeams, ngroups);
    std::vector<char> back = dout.to_host();
    EXPECT(std::memcmp(back.data(), expect.data(), per_beam) == 0);
    // a block too small for the geometry throws
    bool threw = false;
    RawBytes small(blk, per_beam, per_beam);
    try {
        tr(small);
    } catch (std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw);
}

// FengToDada (meerkat/tools/feng_to_dada.cuh): the block is reordered in
// place to the oracle's order before the handler sees it; a block that is
// not a whole number of timestamps throws the reference's error.
void feng_to_dada_reorder()
{
    const std::size_t nchans = 96, nts = 3;
    const std::size_t bytes = nchans * 256 * 2 * sizeof(char2) * nts;
    std::vector<char2> v = gaussian_char2(777, 50.0, bytes / 2, false);
    std::vector<int32_t> words(bytes / 4), expect(bytes / 4);
    std::memcpy(words.data(), v.data(), bytes);
    oracle_feng_to_dada(words.data(), expect.data(), (int)nchans, (long)nts);
    BytesSink sink;
    meerkat::tools::FengToDada<BytesSink> f2d(nchans, sink);
    std::vector<char> hdr(4096, 'h');
    RawBytes hrb(hdr.data(), hdr.size(), hdr.size());
    f2d.init(hrb);
    RawBytes block(reinterpret_cast<char*>(words.data()), bytes, bytes);
    EXPECT(f2d(block) == false);
    EXPECT(sink.calls == 1 && sink.header.size() == 4096);
    EXPECT(std::memcmp(words.data(), expect.data(), bytes) == 0);
    EXPECT(std::memcmp(sink.data.data(), expect.data(), bytes) == 0);
    bool threw = false;
    RawBytes ragged(reinterpret_cast<char*>(words.data()), bytes, bytes - 4);
    try {
        f2d(ragged);
    } catch (std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw && sink.calls == 1);
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
        {"Header.get_set", header_get_set},
        {"DadaRing.write_read_eod", dada_ring_write_read_eod},
        {"DadaRing.cross_process_two_observations", dada_ring_cross_process},
        {"DadaRing.writer_lock_across_processes", dada_ring_writer_lock_across_processes},
        {"DelayManagerTester.test_updates", delay_manager_updates},
        {"DelayManagerTester.test_bad_keys", delay_manager_bad_keys},
        {"PipelineTester.simple_run_test", pipeline_simple_run},
        {"Pipeline.outputs_match_operators_16ant", [] { pipeline_outputs_match_operators(16, 16, 0, 64); }},
        {"Pipeline.outputs_match_operators_64ant_fused", [] { pipeline_outputs_match_operators(64, 64, 0, 256); }},
        {"Pipeline.outputs_match_operators_64of128_taftp",
         [] { pipeline_outputs_match_operators(128, 64, 64, 256); }},
        {"TransposeToDada.beams_match_oracle", tuse_transpose_to_dada},
        {"FengToDada.reorders_in_place", feng_to_dada_reorder},
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
