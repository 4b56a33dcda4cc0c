// The FBFUSE block pipeline (Pipeline.cuh:23-94): a DADA handler that takes
// TAFTP blocks, forms coherent and incoherent beams and writes them to two
// DADA rings.
//
// Same contract as the reference: init(header) reads SAMPLE_CLOCK,
// SAMPLE_CLOCK_START and SYNC_TIME and writes them into both output headers;
// operator()(block) copies block n to the device on call n, processes it on
// call n+1 (delays -> weights at the block's unix time, split transpose,
// coherent beams, incoherent beam) and copies the beams out on call n+2, with
// a/b double buffers between three streams.  The per-block device work is
// enqueued without a host synchronisation after every operator; within one
// call the H2D of block n, the processing of block n-1 and the D2H of the
// beams of block n-2 overlap, and the host waits only where a buffer is
// about to be reused (see operator()).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_PIPELINE_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_PIPELINE_HPP

#include <memory>

#include "psrdada_cpp/dada_write_client.hpp"
#include "psrdada_cpp/double_device_buffer.hpp"
#include "psrdada_cpp/meerkat/fbfuse/DelayManager.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class Pipeline {
public:
    typedef DeviceVector<char2> VoltageVectorType;
    typedef DeviceVector<int8_t> PowerVectorType;
    typedef long double TimeType;

    Pipeline(PipelineConfig const& config, DadaWriteClient& cb_writer, DadaWriteClient& ib_writer,
             std::size_t input_data_buffer_size);
    ~Pipeline();
    Pipeline(Pipeline const&) = delete;

    void init(RawBytes& header);
    bool operator()(RawBytes& data);

    // Unix time of the block processed by the last call (Pipeline.cu:213-215).
    TimeType unix_timestamp() const { return _unix_timestamp; }
    std::size_t call_count() const { return _call_count; }

private:
    void process(VoltageVectorType const& taftp, PowerVectorType& tbtf, PowerVectorType& tf);
    void set_header(RawBytes& header);

    PipelineConfig const& _config;
    std::size_t _sample_clock_start;
    long double _sample_clock;
    long double _sync_time;
    long double _unix_timestamp;
    std::size_t _sample_clock_tick_per_block;
    std::size_t _call_count;
    DoubleDeviceBuffer<char2> _taftp_db;
    DoubleDeviceBuffer<int8_t> _tbtf_db;
    DoubleDeviceBuffer<int8_t> _tf_db;
    DadaWriteClient& _cb_writer;
    DadaWriteClient::HeaderStream& _cb_header_stream;
    DadaWriteClient::DataStream& _cb_data_stream;
    DadaWriteClient& _ib_writer;
    DadaWriteClient::HeaderStream& _ib_header_stream;
    DadaWriteClient::DataStream& _ib_data_stream;
    StreamType _h2d_copy_stream;
    StreamType _processing_stream;
    StreamType _d2h_copy_stream;
    std::size_t _nheap_groups_per_block;
    std::size_t _nsamples_per_dada_block;
    std::unique_ptr<DelayManager> _delay_manager;
    VoltageVectorType _split_transpose_output;
    DeviceVector<char2> _weights;
    fbf_weights_plan* _weights_plan;
    fbf_cb_plan* _cb_plan;
    fbf_ib_plan* _ib_plan;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
