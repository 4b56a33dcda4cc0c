#include "psrdada_cpp/meerkat/fbfuse/IncoherentBeamformer.hpp"

#include <cstring>

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

IncoherentBeamformer::IncoherentBeamformer(PipelineConfig const& config) : _config(config), _plan(nullptr), _desc() {}

IncoherentBeamformer::~IncoherentBeamformer()
{
    if (_plan) fbf_ib_plan_destroy(_plan);
}

fbf_ib_plan* IncoherentBeamformer::plan()
{
    fbf_ib_desc d;
    d.nantennas = (int)_config.ib_nantennas();
    d.nchans = (int)_config.nchans();
    d.nsamples_per_heap = (int)_config.nsamples_per_heap();
    d.tscrunch = (int)_config.ib_tscrunch();
    d.fscrunch = (int)_config.ib_fscrunch();
    d.quant_mode = _config.quantisation_mode();
    std::vector<float> off = _config.ib_power_offsets();
    std::vector<float> sc = _config.ib_power_scalings();
    if (_plan && std::memcmp(&d, &_desc, sizeof(d)) != 0) {
        FBF_ERROR_CHECK(fbf_ib_plan_destroy(_plan));
        _plan = nullptr;
    }
    if (!_plan) {
        FBF_ERROR_CHECK(fbf_ib_plan_create(&d, off.data(), sc.data(), &_plan));
        _desc = d;
        _offsets = off;
        _scalings = sc;
    } else if (off != _offsets || sc != _scalings) {
        FBF_ERROR_CHECK(fbf_ib_plan_set_levels(_plan, off.data(), sc.data()));
        _offsets = off;
        _scalings = sc;
    }
    return _plan;
}

void IncoherentBeamformer::beamform(VoltageVectorType const& input, PowerVectorType& output, StreamType stream)
{
    fbf_ib_plan* p = plan();
    std::size_t out_bytes = 0;
    FBF_ERROR_CHECK(fbf_ib_output_bytes(p, input.bytes(), &out_bytes));
    output.resize(out_bytes);
    FBF_ERROR_CHECK(fbf_ib_beamform(p, input.data(), input.bytes(), output.data(), output.bytes(), stream));
    FBF_ERROR_CHECK(fbf_stream_synchronize(stream));
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
