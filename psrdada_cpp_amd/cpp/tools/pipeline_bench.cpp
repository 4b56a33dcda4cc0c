// Throughput of the drop-in C++ Pipeline (Pipeline.cu:173-247 contract) on
// one DADA block geometry, with the voltages crossing PCIe every block:
// host blocks in pinned memory -> H2D on the copy stream, delays -> weights,
// split-transpose (+ incoherent beam), coherent beams on the processing
// stream, beams D2H into the output rings, whose consumers drop them.
//
//   fbfuse_pipeline_bench [timed_blocks] [warmup_blocks]
//
// Geometry: 64 antennas, 64 channels, 32 heap groups (8192 spectra,
// 39.199 ms), 512 beams, tscrunch 16, the incoherent beam at tscrunch 16.
// Three distinct input blocks (seeded Gaussian voltages) rotate.  Prints one
// JSON line.
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "psrdada_cpp/dada_input_stream.hpp"
#include "psrdada_cpp/dada_write_client.hpp"
#include "psrdada_cpp/meerkat/fbfuse/DelayManager.hpp"
#include "psrdada_cpp/meerkat/fbfuse/Header.hpp"
#include "psrdada_cpp/meerkat/fbfuse/Pipeline.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

using namespace psrdada_cpp;
using namespace psrdada_cpp::meerkat::fbfuse;

int main(int argc, char** argv)
{
    const int timed = argc > 1 ? std::atoi(argv[1]) : 40;
    const int warmup = argc > 2 ? std::atoi(argv[2]) : 8;
    const int nheap_groups = 32;
    const double channel_rate = 1712e6 / (2.0 * 4096);

    PipelineConfig config;
    config.total_nantennas(64);
    config.nchans(64);
    config.cb_nantennas(64);
    config.cb_antenna_offset(0);
    config.cb_nbeams(512);
    config.cb_tscrunch(16);
    config.cb_fscrunch(1);
    config.ib_tscrunch(16);
    config.ib_fscrunch(1);
    config.cb_heap_size(8192);
    config.input_level(16.0f);
    config.output_level(24.0f);
    config.quantisation_mode(FBF_QUANT_ROUND);
    config.centre_frequency(1284e6f);
    config.bandwidth((float)(64 * channel_rate));
    const std::string tag = std::to_string((long)getpid());
    config.delay_buffer_shm("fbfuse_pbench_shm_" + tag);
    config.delay_buffer_mutex("fbfuse_pbench_mutex_" + tag);
    config.delay_buffer_sem("fbfuse_pbench_count_" + tag);

    try {
        DelayEngineSimulator simulator(config);
        // delays tau ~ U(-1 us, 1 us), no rate
        std::vector<float> tau(config.cb_nbeams() * config.cb_nantennas());
        FBF_ERROR_CHECK(fbf_synth_uniform_f32(2, -1e-6f, 1e-6f, tau.data(), tau.size()));
        float2* d = simulator.delays();
        for (std::size_t i = 0; i < tau.size(); ++i) d[i] = float2{0.0f, tau[i]};
        simulator.update_delays();

        const std::size_t nsamples = (std::size_t)nheap_groups * config.nsamples_per_heap();
        const std::size_t in_bytes = (std::size_t)nheap_groups * config.total_nantennas() * config.nchans() *
                                     config.nsamples_per_heap() * config.npol() * sizeof(char2);
        const std::size_t cb_bytes = config.cb_nbeams() * (nsamples / config.cb_tscrunch()) * config.nchans();
        const std::size_t ib_bytes = (nsamples / config.ib_tscrunch()) * config.nchans();
        // Output rings in this process with pinned blocks (the reference pins
        // its rings for the device copies, fbfuse_cli.cu:133-135).
        DadaDB cb_db(8, cb_bytes, 4, 4096, DadaRing::Placement::process);
        DadaDB ib_db(8, ib_bytes, 4, 4096, DadaRing::Placement::process);
        cb_db.create();
        ib_db.create();
        config.cb_dada_key(cb_db.key());
        config.ib_dada_key(ib_db.key());
        MultiLog log("pipeline_bench");
        DadaWriteClient cb_writer(config.cb_dada_key(), log), ib_writer(config.ib_dada_key(), log);

        const int nin = 3;
        std::vector<void*> inputs(nin, nullptr);
        for (int i = 0; i < nin; ++i) {
            FBF_ERROR_CHECK(fbf_host_malloc(&inputs[i], in_bytes));
            FBF_ERROR_CHECK(fbf_synth_gaussian_int8(20260101 + i, 16.0, 1, static_cast<int8_t*>(inputs[i]), in_bytes));
        }
        std::vector<char> header_buffer(4096, 0);
        RawBytes header_rb(header_buffer.data(), header_buffer.size(), header_buffer.size());
        Header header(header_rb);
        header.set<long double>("SAMPLE_CLOCK", 1712000000.0L);
        header.set<long double>("SYNC_TIME", 1700000000.0L);
        header.set<std::size_t>("SAMPLE_CLOCK_START", 0);

        double ms_per_block = 0.0;
        {
            Pipeline pipeline(config, cb_writer, ib_writer, in_bytes);
            NullSink sink;
            DadaInputStream<NullSink> cb_consumer(config.cb_dada_key(), log, sink);
            DadaInputStream<NullSink> ib_consumer(config.ib_dada_key(), log, sink);
            std::thread cb_thread([&] { cb_consumer.start(); });
            std::thread ib_thread([&] { ib_consumer.start(); });
            pipeline.init(header_rb);
            int call = 0;
            auto block = [&]() {
                RawBytes rb(static_cast<char*>(inputs[call % nin]), in_bytes, in_bytes);
                pipeline(rb);
                ++call;
            };
            for (int i = 0; i < warmup; ++i) block();
            const auto t0 = std::chrono::steady_clock::now();
            for (int i = 0; i < timed; ++i) block();
            const auto t1 = std::chrono::steady_clock::now();
            ms_per_block = std::chrono::duration<double, std::milli>(t1 - t0).count() / timed;
            cb_consumer.stop();
            ib_consumer.stop();
            block();   // wakes the consumers
            block();
            cb_thread.join();
            ib_thread.join();
        }
        for (void* p : inputs) fbf_host_free(p);
        const double block_ms = nsamples / channel_rate * 1e3;
        const double ops = 8.0 * config.nchans() * nsamples * 2 * config.cb_nantennas() * config.cb_nbeams();
        std::printf("{\"what\": \"C++ Pipeline (Pipeline.cu:173-247 contract): pinned host block -> H2D; coherent "
                    "and incoherent beams in one pass over TAFTP, the next block's weights in the reorder's "
                    "launch; D2H into the output rings; three streams\", "
                    "\"geometry\": \"64 ant, 64 ch, 32 heap groups, 512 beams, ts 16, IB ts 16\", "
                    "\"input\": \"3 rotating pinned host blocks\", \"blocks\": %d, \"warmup\": %d, "
                    "\"ms_per_block\": %.4f, \"x_real_time\": %.3f, \"h2d_gbs\": %.2f, \"int8_tops\": %.2f, "
                    "\"block_ms\": %.3f}\n",
                    timed, warmup, ms_per_block, block_ms / ms_per_block, in_bytes / (ms_per_block * 1e-3) / 1e9,
                    ops / (ms_per_block * 1e-3) / 1e12, block_ms);
    } catch (std::exception const& e) {
        std::fprintf(stderr, "pipeline_bench: %s\n", e.what());
        return 1;
    }
    return 0;
}
