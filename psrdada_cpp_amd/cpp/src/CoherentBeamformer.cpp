#include "psrdada_cpp/meerkat/fbfuse/CoherentBeamformer.hpp"

#include <cstring>

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

CoherentBeamformer::CoherentBeamformer(PipelineConfig const& config)
    : _config(config), _plan(nullptr), _desc(), _variant(0)
{
}

CoherentBeamformer::~CoherentBeamformer()
{
    if (_plan) fbf_cb_plan_destroy(_plan);
}

// The plan follows the configuration: geometry changes rebuild it, level
// changes are pushed to it (the reference reads both from the config on
// every call, CoherentBeamformer.cu:210-222).
fbf_cb_plan* CoherentBeamformer::plan()
{
    fbf_cb_desc d;
    d.nantennas = (int)_config.cb_nantennas();
    d.nbeams = (int)_config.cb_nbeams();
    d.nchans = (int)_config.nchans();
    d.tscrunch = (int)_config.cb_tscrunch();
    d.fscrunch = (int)_config.cb_fscrunch();
    d.nsamples_per_heap_out = (int)_config.cb_nsamples_per_heap();
    d.quant_mode = _config.quantisation_mode();
    std::vector<float> off = _config.cb_power_offsets();
    std::vector<float> sc = _config.cb_power_scalings();
    if (_plan && std::memcmp(&d, &_desc, sizeof(d)) != 0) {
        FBF_ERROR_CHECK(fbf_cb_plan_destroy(_plan));
        _plan = nullptr;
    }
    if (!_plan) {
        FBF_ERROR_CHECK(fbf_cb_plan_create(&d, off.data(), sc.data(), &_plan));
        FBF_ERROR_CHECK(fbf_cb_plan_set_variant(_plan, _variant));
        _desc = d;
        _offsets = off;
        _scalings = sc;
    } else if (off != _offsets || sc != _scalings) {
        FBF_ERROR_CHECK(fbf_cb_plan_set_levels(_plan, off.data(), sc.data()));
        _offsets = off;
        _scalings = sc;
    }
    return _plan;
}

void CoherentBeamformer::variant(int v)
{
    _variant = v;
    if (_plan) FBF_ERROR_CHECK(fbf_cb_plan_set_variant(_plan, v));
}

void CoherentBeamformer::beamform(VoltageVectorType const& input, WeightsVectorType const& weights,
                                  PowerVectorType& output, StreamType stream)
{
    fbf_cb_plan* p = plan();
    std::size_t out_bytes = 0;
    FBF_ERROR_CHECK(fbf_cb_output_bytes(p, input.bytes(), &out_bytes));
    output.resize(out_bytes);
    FBF_ERROR_CHECK(fbf_cb_beamform(p, input.data(), input.bytes(), weights.data(), weights.bytes(), output.data(),
                                    output.bytes(), stream));
    FBF_ERROR_CHECK(fbf_stream_synchronize(stream));
}

void CoherentBeamformer::beam_power(VoltageVectorType const& input, WeightsVectorType const& weights,
                                    DeviceVector<double>& output, StreamType stream)
{
    fbf_cb_plan* p = plan();
    std::size_t out_bytes = 0;
    FBF_ERROR_CHECK(fbf_cb_output_bytes(p, input.bytes(), &out_bytes));
    output.resize(out_bytes);
    FBF_ERROR_CHECK(fbf_cb_beam_power(p, input.data(), input.bytes(), weights.data(), weights.bytes(), output.data(),
                                      output.bytes(), stream));
    FBF_ERROR_CHECK(fbf_stream_synchronize(stream));
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
