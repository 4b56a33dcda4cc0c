// Incoherent beamformer operator (IncoherentBeamformer.cuh:27-59).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_INCOHERENTBEAMFORMER_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_INCOHERENTBEAMFORMER_HPP

#include <vector>

#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class IncoherentBeamformer {
public:
    typedef DeviceVector<char2> VoltageVectorType;   // TAFTP
    typedef DeviceVector<int8_t> PowerVectorType;    // TF

    explicit IncoherentBeamformer(PipelineConfig const& config);
    ~IncoherentBeamformer();
    IncoherentBeamformer(IncoherentBeamformer const&) = delete;

    // TAFTP voltages -> int8 TF powers (all antennas).  The output is resized
    // here; the stream is synchronised before returning
    // (IncoherentBeamformer.cu:89-117).
    void beamform(VoltageVectorType const& input, PowerVectorType& output, StreamType stream);

private:
    fbf_ib_plan* plan();

    PipelineConfig const& _config;
    fbf_ib_plan* _plan;
    fbf_ib_desc _desc;
    std::vector<float> _offsets, _scalings;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
