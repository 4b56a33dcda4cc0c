// Coherent beamformer operator (CoherentBeamformer.cuh:100-137).
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_COHERENTBEAMFORMER_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_COHERENTBEAMFORMER_HPP

#include <vector>

#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/meerkat/fbfuse/PipelineConfig.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class CoherentBeamformer {
public:
    typedef DeviceVector<char2> VoltageVectorType;   // FTPA
    typedef DeviceVector<int8_t> PowerVectorType;    // TBTF
    typedef DeviceVector<char2> WeightsVectorType;   // FBA

    explicit CoherentBeamformer(PipelineConfig const& config);
    ~CoherentBeamformer();
    CoherentBeamformer(CoherentBeamformer const&) = delete;

    // FTPA voltages x FBA weights -> int8 TBTF powers.  The output is resized
    // here; the stream is synchronised before returning
    // (CoherentBeamformer.cu:190-225).  Size errors throw.
    void beamform(VoltageVectorType const& input, WeightsVectorType const& weights, PowerVectorType& output,
                  StreamType stream);

    // The exact detected power before quantisation, double [F/fs][T/ts][B].
    void beam_power(VoltageVectorType const& input, WeightsVectorType const& weights, DeviceVector<double>& output,
                    StreamType stream);

    // Kernel variant (0 = automatic); see fbf_cb_variant_count.
    void variant(int v);

private:
    fbf_cb_plan* plan();

    PipelineConfig const& _config;
    fbf_cb_plan* _plan;
    fbf_cb_desc _desc;
    std::vector<float> _offsets, _scalings;
    int _variant;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
