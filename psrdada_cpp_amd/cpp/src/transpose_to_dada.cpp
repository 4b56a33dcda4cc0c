#include "psrdada_cpp/meerkat/tuse/transpose_to_dada.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace tuse {
namespace transpose {

void do_transpose(RawBytes& transposed_data, RawBytes& input_data, std::uint32_t nchans, std::uint32_t nsamples,
                  std::uint32_t nfreq, std::uint32_t beamnum, std::uint32_t nbeams, std::uint32_t ngroups)
{
    const std::size_t per_beam = (std::size_t)nchans * nsamples * nfreq;
    const std::size_t need_in = per_beam * nbeams * ngroups;
    const std::size_t need_out = per_beam * ngroups;
    if (input_data.total_bytes() < need_in)
        throw std::runtime_error("do_transpose: input holds " + std::to_string(input_data.total_bytes()) +
                                 " bytes, the geometry needs " + std::to_string(need_in));
    if (transposed_data.total_bytes() < need_out)
        throw std::runtime_error("do_transpose: output holds " + std::to_string(transposed_data.total_bytes()) +
                                 " bytes, one beam needs " + std::to_string(need_out));
    if (input_data.on_device() != transposed_data.on_device())
        throw std::runtime_error("do_transpose: input and output must both be on the host or both on the device");
    if (input_data.on_device()) {
        FBF_ERROR_CHECK(fbf_tuse_transpose(input_data.ptr(), input_data.total_bytes(), transposed_data.ptr(),
                                           need_out, (int)nchans, (int)nsamples, (int)nfreq, (int)nbeams,
                                           (int)ngroups, (int)beamnum, 1, nullptr));
        FBF_ERROR_CHECK(fbf_stream_synchronize(nullptr));
        return;
    }
    DeviceVector<char> in(need_in), out(need_out);
    FBF_ERROR_CHECK(fbf_memcpy_h2d(in.data(), input_data.ptr(), need_in, nullptr));
    FBF_ERROR_CHECK(fbf_tuse_transpose(in.data(), need_in, out.data(), need_out, (int)nchans, (int)nsamples,
                                       (int)nfreq, (int)nbeams, (int)ngroups, (int)beamnum, 1, nullptr));
    FBF_ERROR_CHECK(fbf_memcpy_d2h(transposed_data.ptr(), out.data(), need_out, nullptr));
    FBF_ERROR_CHECK(fbf_stream_synchronize(nullptr));
}

}  // namespace transpose
}  // namespace tuse
}  // namespace meerkat
}  // namespace psrdada_cpp
