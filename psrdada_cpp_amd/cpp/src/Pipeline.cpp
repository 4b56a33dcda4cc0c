#include "psrdada_cpp/meerkat/fbfuse/Pipeline.hpp"

#include <stdexcept>
#include <string>

#include "psrdada_cpp/meerkat/fbfuse/Header.hpp"

#define FBFUSE_SAMPLE_CLOCK_START_KEY "SAMPLE_CLOCK_START"
#define FBFUSE_SAMPLE_CLOCK_KEY "SAMPLE_CLOCK"
#define FBFUSE_SYNC_TIME_KEY "SYNC_TIME"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

Pipeline::Pipeline(PipelineConfig const& config, DadaWriteClient& cb_writer, DadaWriteClient& ib_writer,
                   std::size_t input_data_buffer_size)
    : _config(config),
      _sample_clock_start(0),
      _sample_clock(0.0),
      _sync_time(0.0),
      _unix_timestamp(0.0),
      _sample_clock_tick_per_block(0),
      _call_count(0),
      _cb_writer(cb_writer),
      _cb_header_stream(cb_writer.header_stream()),
      _cb_data_stream(cb_writer.data_stream()),
      _ib_writer(ib_writer),
      _ib_header_stream(ib_writer.header_stream()),
      _ib_data_stream(ib_writer.data_stream()),
      _h2d_copy_stream(nullptr),
      _processing_stream(nullptr),
      _d2h_copy_stream(nullptr),
      _weights_plan(nullptr),
      _cb_plan(nullptr),
      _ib_plan(nullptr)
{
    // Input and output buffer checks (Pipeline.cu:42-88).
    const std::size_t heap_group_size =
        _config.total_nantennas() * _config.nchans() * _config.nsamples_per_heap() * _config.npol() * sizeof(char2);
    if (input_data_buffer_size % heap_group_size != 0)
        throw std::runtime_error("Pipeline: input blocks of " + std::to_string(input_data_buffer_size) +
                                 " bytes are not a whole number of " + std::to_string(heap_group_size) +
                                 "-byte heap groups");
    _nheap_groups_per_block = input_data_buffer_size / heap_group_size;
    _nsamples_per_dada_block = _nheap_groups_per_block * _config.nsamples_per_heap();
    if (_nsamples_per_dada_block % _config.cb_tscrunch() != 0 ||
        (_nsamples_per_dada_block / _config.cb_tscrunch()) % _config.cb_nsamples_per_heap() != 0)
        throw std::runtime_error("Pipeline: the spectra of one input block do not fill whole coherent-beam "
                                 "output heaps");
    _taftp_db.resize(input_data_buffer_size / sizeof(char2), {0, 0});

    const std::size_t expected_cb_size = (_config.cb_nbeams() * _nsamples_per_dada_block / _config.cb_tscrunch() *
                                          _config.nchans() / _config.cb_fscrunch());
    if (_cb_writer.data_buffer_size() != expected_cb_size)
        throw std::runtime_error("Pipeline: one input block gives " + std::to_string(expected_cb_size) +
                                 " bytes of coherent beams, but the output ring's blocks hold " +
                                 std::to_string(_cb_writer.data_buffer_size()));
    _tbtf_db.resize(expected_cb_size, 0);
    const std::size_t expected_ib_size =
        (_config.ib_nbeams() * _nsamples_per_dada_block / _config.ib_tscrunch() * _config.nchans() /
         _config.ib_fscrunch());
    if (_ib_writer.data_buffer_size() != expected_ib_size)
        throw std::runtime_error("Pipeline: one input block gives " + std::to_string(expected_ib_size) +
                                 " bytes of incoherent beam, but the output ring's blocks hold " +
                                 std::to_string(_ib_writer.data_buffer_size()));
    _tf_db.resize(expected_ib_size, 0);

    _sample_clock_tick_per_block = 2 * _config.total_nchans() * _nsamples_per_dada_block;   // Pipeline.cu:91
    FBF_ERROR_CHECK(fbf_stream_create(&_h2d_copy_stream));
    FBF_ERROR_CHECK(fbf_stream_create(&_processing_stream));
    FBF_ERROR_CHECK(fbf_stream_create(&_d2h_copy_stream));

    _delay_manager.reset(new DelayManager(_config, _h2d_copy_stream));
    _weights.resize(_config.cb_nbeams() * _config.cb_nantennas() * _config.nchans());
    std::vector<double> freqs = _config.channel_frequencies_f64();
    FBF_ERROR_CHECK(fbf_weights_plan_create((int)_config.cb_nantennas(), (int)_config.cb_nbeams(),
                                            (int)_config.nchans(), freqs.data(), &_weights_plan));
    fbf_cb_desc cbd{(int)_config.cb_nantennas(), (int)_config.cb_nbeams(), (int)_config.nchans(),
                    (int)_config.cb_tscrunch(), (int)_config.cb_fscrunch(), (int)_config.cb_nsamples_per_heap(),
                    _config.quantisation_mode()};
    std::vector<float> cbo = _config.cb_power_offsets(), cbs = _config.cb_power_scalings();
    FBF_ERROR_CHECK(fbf_cb_plan_create(&cbd, cbo.data(), cbs.data(), &_cb_plan));
    fbf_ib_desc ibd{(int)_config.ib_nantennas(), (int)_config.nchans(), (int)_config.nsamples_per_heap(),
                    (int)_config.ib_tscrunch(), (int)_config.ib_fscrunch(), _config.quantisation_mode()};
    std::vector<float> ibo = _config.ib_power_offsets(), ibs = _config.ib_power_scalings();
    FBF_ERROR_CHECK(fbf_ib_plan_create(&ibd, ibo.data(), ibs.data(), &_ib_plan));
}

Pipeline::~Pipeline()
{
    try {
        if (_call_count > 2) {
            fbf_stream_synchronize(_d2h_copy_stream);
            _cb_data_stream.release();
            _ib_data_stream.release();
        }
    } catch (std::exception const&) {
        // non-fatal on destruction (Pipeline.cu:109-118)
    }
    fbf_stream_synchronize(_processing_stream);
    fbf_stream_synchronize(_h2d_copy_stream);
    if (_weights_plan) fbf_weights_plan_destroy(_weights_plan);
    if (_cb_plan) fbf_cb_plan_destroy(_cb_plan);
    if (_ib_plan) fbf_ib_plan_destroy(_ib_plan);
    fbf_stream_destroy(_h2d_copy_stream);
    fbf_stream_destroy(_processing_stream);
    fbf_stream_destroy(_d2h_copy_stream);
}

void Pipeline::set_header(RawBytes& header)
{
    Header parser(header);
    parser.purge();
    parser.set<long double>(FBFUSE_SAMPLE_CLOCK_KEY, _sample_clock);
    parser.set<long double>(FBFUSE_SYNC_TIME_KEY, _sync_time);
    parser.set<std::size_t>(FBFUSE_SAMPLE_CLOCK_START_KEY, _sample_clock_start);
    header.used_bytes(header.total_bytes());
}

void Pipeline::init(RawBytes& header)
{
    Header parser(header);
    _sample_clock_start = parser.get<std::size_t>(FBFUSE_SAMPLE_CLOCK_START_KEY);
    _sample_clock = parser.get<long double>(FBFUSE_SAMPLE_CLOCK_KEY);
    _sync_time = parser.get<long double>(FBFUSE_SYNC_TIME_KEY);
    auto& cb_header_block = _cb_header_stream.next();
    set_header(cb_header_block);
    _cb_header_stream.release();
    auto& ib_header_block = _ib_header_stream.next();
    set_header(ib_header_block);
    _ib_header_stream.release();
}

void Pipeline::process(VoltageVectorType const& taftp, PowerVectorType& tbtf, PowerVectorType& tf)
{
    auto const& delays = _delay_manager->delays();
    int one_pass = 0, direct = 0;
    if (_config.cb_nantennas() == _config.total_nantennas() && _config.cb_antenna_offset() == 0)
        FBF_ERROR_CHECK(fbf_cb_ib_fused(_cb_plan, _ib_plan, (int)_config.nsamples_per_heap(),
                                        (long long)_nsamples_per_dada_block, &one_pass));
    if (one_pass) {
        // Both beams in one pass over the TAFTP block (the coherent kernel
        // leaves the incoherent beam's partials), and the reorder's launch
        // generates the next block's weights.  This block's weights came from
        // the previous block's launch when they were made for this epoch with
        // the delay model still in force; else they are generated first.  The
        // outputs are those of the reference's order (Pipeline.cu:164-170).
        const double epoch = (double)_unix_timestamp;
        if (!(_next_weights_valid && _next_weights_epoch == epoch &&
              _next_weights_updates == _delay_manager->update_count()))
            FBF_ERROR_CHECK(fbf_weights_generate(_weights_plan, delays.data(), delays.bytes(), epoch,
                                                 _weights.data(), _weights.bytes(), nullptr, _processing_stream));
        const double next_epoch =
            (double)(_sync_time + (_sample_clock_start + ((_call_count - 1) * _sample_clock_tick_per_block)) /
                                      _sample_clock);
        FBF_ERROR_CHECK(fbf_cb_ib_beamform_taftp_next(
            _cb_plan, _ib_plan, taftp.data(), taftp.bytes(), (int)_config.nsamples_per_heap(), _weights.data(),
            _weights.bytes(), tbtf.data(), tbtf.bytes(), tf.data(), tf.bytes(), _weights_plan, delays.data(),
            delays.bytes(), next_epoch, _weights.data(), _weights.bytes(), _processing_stream));
        _next_weights_valid = true;
        _next_weights_epoch = next_epoch;
        _next_weights_updates = _delay_manager->update_count();
        return;
    }
    _next_weights_valid = false;
    FBF_ERROR_CHECK(fbf_cb_taftp_direct(_cb_plan, (int)_config.total_nantennas(), (int)_config.nsamples_per_heap(),
                                        (long long)_nsamples_per_dada_block, &direct));
    if (direct) {
        // The weights in the incoherent beam's launch, then the coherent
        // beams straight from the TAFTP block: no split-transpose pass, the
        // same outputs as the reference's order (Pipeline.cu:164-170).
        FBF_ERROR_CHECK(fbf_ib_beamform_weights(_ib_plan, taftp.data(), taftp.bytes(), tf.data(), tf.bytes(),
                                                _weights_plan, delays.data(), delays.bytes(),
                                                (double)_unix_timestamp, _weights.data(), _weights.bytes(),
                                                _processing_stream));
        FBF_ERROR_CHECK(fbf_cb_beamform_taftp(_cb_plan, taftp.data(), taftp.bytes(), (int)_config.total_nantennas(),
                                              (int)_config.cb_antenna_offset(), (int)_config.nsamples_per_heap(),
                                              _weights.data(), _weights.bytes(), tbtf.data(), tbtf.bytes(),
                                              _processing_stream));
        return;
    }
    FBF_ERROR_CHECK(fbf_weights_generate(_weights_plan, delays.data(), delays.bytes(), (double)_unix_timestamp,
                                         _weights.data(), _weights.bytes(), nullptr, _processing_stream));
    _split_transpose_output.resize(_nsamples_per_dada_block * _config.npol() * _config.nchans() *
                                   _config.cb_nantennas());
    // A whole-block split can produce the incoherent beam in the same pass
    // over the input when the IB scrunch allows it.
    const bool fused = _config.cb_nantennas() == _config.total_nantennas() && _config.cb_antenna_offset() == 0 &&
                       _config.ib_fscrunch() == 1 && 32 % _config.ib_tscrunch() == 0 &&
                       _config.nsamples_per_heap() % 32 == 0;
    if (fused) {
        FBF_ERROR_CHECK(fbf_split_transpose_ib(taftp.data(), taftp.bytes(), _split_transpose_output.data(),
                                               _split_transpose_output.bytes(), (int)_config.nchans(),
                                               (int)_config.nsamples_per_heap(), _ib_plan, tf.data(), tf.bytes(),
                                               _processing_stream));
    } else {
        FBF_ERROR_CHECK(fbf_split_transpose(taftp.data(), taftp.bytes(), _split_transpose_output.data(),
                                            _split_transpose_output.bytes(), (int)_config.total_nantennas(),
                                            (int)_config.cb_nantennas(), (int)_config.cb_antenna_offset(),
                                            (int)_config.nchans(), (int)_config.nsamples_per_heap(),
                                            _processing_stream));
    }
    FBF_ERROR_CHECK(fbf_cb_beamform(_cb_plan, _split_transpose_output.data(), _split_transpose_output.bytes(),
                                    _weights.data(), _weights.bytes(), tbtf.data(), tbtf.bytes(),
                                    _processing_stream));
    if (!fused)
        FBF_ERROR_CHECK(
            fbf_ib_beamform(_ib_plan, taftp.data(), taftp.bytes(), tf.data(), tf.bytes(), _processing_stream));
}

bool Pipeline::operator()(RawBytes& data)
{
    // Per call n: H2D of block n, processing of block n-1 and D2H of the
    // beams of block n-2 run concurrently on their three streams.  The
    // synchronisation points below make every buffer reuse safe: the
    // processing of block n-1 and the D2H of block n-2 finished before their
    // buffers are swapped and written again, and the H2D of block n finished
    // before this call returns (the caller releases `data` afterwards).
    ++_call_count;
    FBF_ERROR_CHECK(fbf_stream_synchronize(_processing_stream));
    FBF_ERROR_CHECK(fbf_stream_synchronize(_d2h_copy_stream));
    _taftp_db.swap();
    if (data.used_bytes() != _taftp_db.a().size() * sizeof(char2))
        throw std::runtime_error("Pipeline: a block of " + std::to_string(data.used_bytes()) +
                                 " bytes arrived; the pipeline was set up for " +
                                 std::to_string(_taftp_db.a().size() * sizeof(char2)));
    FBF_ERROR_CHECK(fbf_memcpy_h2d(_taftp_db.a_ptr(), data.ptr(), data.used_bytes(), _h2d_copy_stream));
    if (_call_count == 1) {
        FBF_ERROR_CHECK(fbf_stream_synchronize(_h2d_copy_stream));
        return false;
    }

    _tbtf_db.swap();
    _tf_db.swap();
    // Unix time of the block processed now (Pipeline.cu:213-215).
    _unix_timestamp =
        (_sync_time + (_sample_clock_start + ((_call_count - 2) * _sample_clock_tick_per_block)) / _sample_clock);
    process(_taftp_db.b(), _tbtf_db.a(), _tf_db.a());
    if (_call_count == 2) {
        FBF_ERROR_CHECK(fbf_stream_synchronize(_h2d_copy_stream));
        return false;
    }

    if (_call_count > 3) {
        _cb_data_stream.release();
        _ib_data_stream.release();
    }
    auto& cb_block = _cb_data_stream.next();
    auto& ib_block = _ib_data_stream.next();
    FBF_ERROR_CHECK(fbf_memcpy_d2h(cb_block.ptr(), _tbtf_db.b_ptr(), cb_block.total_bytes(), _d2h_copy_stream));
    FBF_ERROR_CHECK(fbf_memcpy_d2h(ib_block.ptr(), _tf_db.b_ptr(), ib_block.total_bytes(), _d2h_copy_stream));
    cb_block.used_bytes(cb_block.total_bytes());
    ib_block.used_bytes(ib_block.total_bytes());
    FBF_ERROR_CHECK(fbf_stream_synchronize(_h2d_copy_stream));
    return false;
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
