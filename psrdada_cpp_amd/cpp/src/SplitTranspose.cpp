#include "psrdada_cpp/meerkat/fbfuse/SplitTranspose.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

SplitTranspose::SplitTranspose(PipelineConfig const& config) : _config(config) {}

SplitTranspose::~SplitTranspose() {}

void SplitTranspose::transpose(VoltageType const& taftp_voltages, VoltageType& ftpa_voltages, StreamType stream)
{
    const std::size_t heap_group_size =
        _config.npol() * _config.nsamples_per_heap() * _config.nchans() * _config.total_nantennas();
    if (taftp_voltages.size() % heap_group_size != 0)
        throw std::runtime_error("SplitTranspose: input is not a whole number of heap groups");
    const std::size_t nheap_groups = taftp_voltages.size() / heap_group_size;
    const std::size_t out_per_hg =
        _config.npol() * _config.nsamples_per_heap() * _config.nchans() * _config.cb_nantennas();
    ftpa_voltages.resize(out_per_hg * nheap_groups);
    FBF_ERROR_CHECK(fbf_split_transpose(taftp_voltages.data(), taftp_voltages.bytes(), ftpa_voltages.data(),
                                        ftpa_voltages.bytes(), (int)_config.total_nantennas(),
                                        (int)_config.cb_nantennas(), (int)_config.cb_antenna_offset(),
                                        (int)_config.nchans(), (int)_config.nsamples_per_heap(), stream));
    FBF_ERROR_CHECK(fbf_stream_synchronize(stream));
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp
