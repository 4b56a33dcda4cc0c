// TAFTP -> FTPA split transpose (SplitTranspose.cuh:51-81).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_SPLITTRANSPOSE_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_SPLITTRANSPOSE_HPP

#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class SplitTranspose {
public:
    typedef DeviceVector<char2> VoltageType;

    explicit SplitTranspose(PipelineConfig const& config);
    ~SplitTranspose();
    SplitTranspose(SplitTranspose const&) = delete;

    // Selects antennas [cb_antenna_offset, cb_antenna_offset + cb_nantennas)
    // and reorders TAFTP -> FTPA.  The output is resized here and the stream
    // is synchronised before returning, as in SplitTranspose.cu:98-126.
    void transpose(VoltageType const& taftp_voltages, VoltageType& ftpa_voltages, StreamType stream);

private:
    PipelineConfig const& _config;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
