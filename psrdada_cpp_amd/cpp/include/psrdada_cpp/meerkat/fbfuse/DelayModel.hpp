// Layout of the delay-model shared memory written by the control system
// (DelayModel.cuh:11-16): an FP64 epoch and validity duration followed by one
// float2 (x = rate s/s, y = offset s) per (beam, antenna), beams slowest.
// The reference sizes the array at compile time; here it is
// nbeams * nantennas of the configuration, with the same byte layout.
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_DELAYMODEL_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_DELAYMODEL_HPP

#include <cstddef>

#include "psrdada_cpp/common.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

struct DelayModelHeader {
    double epoch;
    double duration;
};

inline std::size_t delay_model_bytes(std::size_t nbeams, std::size_t nantennas)
{
    return sizeof(DelayModelHeader) + nbeams * nantennas * sizeof(float2);
}

inline float2* delay_model_delays(void* model)
{
    return reinterpret_cast<float2*>(static_cast<char*>(model) + sizeof(DelayModelHeader));
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
