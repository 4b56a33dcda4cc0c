// Delay model -> beamforming weights (WeightsManager.cuh:27-77).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_WEIGHTSMANAGER_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_WEIGHTSMANAGER_HPP

#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class WeightsManager {
public:
    typedef char2 WeightsType;
    typedef DeviceVector<WeightsType> WeightsVectorType;
    typedef float FreqType;
    // The reference takes the epoch as float (WeightsManager.cuh:34); the
    // phase argument here is formed in FP64, so the epoch is a double (a
    // float argument converts implicitly).
    typedef double TimeType;
    typedef DeviceVector<float2> DelayVectorType;

    WeightsManager(PipelineConfig const& config, StreamType stream);
    ~WeightsManager();
    WeightsManager(WeightsManager const&) = delete;

    // Weights [F][B][A] for the delay models [B][A] (x = rate s/s, y = offset
    // s) evaluated at `epoch` seconds.  Returns a reference to an internal
    // buffer that the next call overwrites (not thread-safe, as in the
    // reference).  The stream is synchronised before returning.
    WeightsVectorType const& weights(DelayVectorType const& delays, TimeType epoch);

private:
    PipelineConfig const& _config;
    StreamType _stream;
    WeightsVectorType _weights;
    fbf_weights_plan* _plan;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
