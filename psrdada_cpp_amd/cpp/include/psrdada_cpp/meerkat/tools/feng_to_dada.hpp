// F-engine heaps to DADA order (meerkat/tools/feng_to_dada.cuh:21-64,
// meerkat/tools/detail/feng_to_dada.cu:9-52).
//
// The same names, constructor and error as the reference: each block is a
// whole number of timestamps of [nchans][256 samples][2 pols] char2 heaps;
// the reorder to [256][nchans][2 pols] runs on the GPU (fbf_feng_to_dada)
// and the result is copied back into the block before the handler sees it.
#ifndef PSRDADA_CPP_MEERKAT_TOOLS_FENG_TO_DADA_HPP
#define PSRDADA_CPP_MEERKAT_TOOLS_FENG_TO_DADA_HPP

#include <cstddef>

#include "psrdada_cpp/common.hpp"
#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace tools {

// meerkat/constants.hpp:4-6
constexpr std::size_t kFengNsampsPerHeap = 256;
constexpr std::size_t kFengNpolPerHeap = 2;

template <class HandlerType>
class FengToDada {
public:
    FengToDada(std::size_t nchans, HandlerType& handler) : _nchans(nchans), _handler(handler) {}
    ~FengToDada() {}
    FengToDada(FengToDada const&) = delete;

    void init(RawBytes& block) { _handler.init(block); }

    bool operator()(RawBytes& block)
    {
        const std::size_t used = block.used_bytes();
        const std::size_t nbytes_per_timestamp = _nchans * kFengNsampsPerHeap * kFengNpolPerHeap * sizeof(char2);
        if (used % nbytes_per_timestamp != 0)
            throw std::runtime_error(
                "Number of bytes in buffer is not an integer muliple of the number of bytes per timestamp");
        if (used) {
            _input.resize(used);
            _output.resize(used);
            if (block.on_device())
                FBF_ERROR_CHECK(fbf_memcpy_d2d(_input.data(), block.ptr(), used, nullptr));
            else
                FBF_ERROR_CHECK(fbf_memcpy_h2d(_input.data(), block.ptr(), used, nullptr));
            FBF_ERROR_CHECK(fbf_feng_to_dada(_input.data(), _output.data(), used, (int)_nchans, nullptr));
            if (block.on_device())
                FBF_ERROR_CHECK(fbf_memcpy_d2d(block.ptr(), _output.data(), used, nullptr));
            else
                FBF_ERROR_CHECK(fbf_memcpy_d2h(block.ptr(), _output.data(), used, nullptr));
            FBF_ERROR_CHECK(fbf_stream_synchronize(nullptr));
        }
        _handler(block);
        return false;
    }

private:
    std::size_t _nchans;
    HandlerType& _handler;
    DeviceVector<char> _input;
    DeviceVector<char> _output;
};

}  // namespace tools
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
