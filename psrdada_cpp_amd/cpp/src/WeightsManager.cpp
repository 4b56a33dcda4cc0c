#include "psrdada_cpp/meerkat/fbfuse/WeightsManager.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

WeightsManager::WeightsManager(PipelineConfig const& config, StreamType stream)
    : _config(config), _stream(stream), _plan(nullptr)
{
    _weights.resize(_config.cb_nbeams() * _config.cb_nantennas() * _config.nchans());
    std::vector<double> freqs = _config.channel_frequencies_f64();
    FBF_ERROR_CHECK(fbf_weights_plan_create((int)_config.cb_nantennas(), (int)_config.cb_nbeams(),
                                            (int)_config.nchans(), freqs.data(), &_plan));
}

WeightsManager::~WeightsManager()
{
    if (_plan) fbf_weights_plan_destroy(_plan);
}

WeightsManager::WeightsVectorType const& WeightsManager::weights(DelayVectorType const& delays, TimeType epoch)
{
    FBF_ERROR_CHECK(fbf_weights_generate(_plan, delays.data(), delays.bytes(), epoch, _weights.data(),
                                         _weights.bytes(), nullptr, _stream));
    FBF_ERROR_CHECK(fbf_stream_synchronize(_stream));
    return _weights;
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
