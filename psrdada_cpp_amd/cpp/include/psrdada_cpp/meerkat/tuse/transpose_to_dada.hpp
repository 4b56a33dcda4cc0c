// TUSE beam corner turn (meerkat/tuse/transpose_to_dada.hpp:11-121).
//
// The same names, setters, getters and defaults as the reference.  The
// corner turn runs on the GPU (fbf_tuse_transpose): one host->device copy of
// the block, one launch for every beam, one copy back, then each beam's
// handler is called (in its own thread, as in the reference) with that
// beam's own buffer -- the reference had every thread write one shared
// buffer.  GPU and size errors throw std::runtime_error.
#ifndef PSRDADA_CPP_MEERKAT_TUSE_TRANSPOSE_TO_DADA_HPP
#define PSRDADA_CPP_MEERKAT_TUSE_TRANSPOSE_TO_DADA_HPP

#include <cstdint>
#include <memory>
#include <thread>
#include <vector>

#include "psrdada_cpp/common.hpp"
#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/raw_bytes.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace tuse {

namespace transpose {

// Writes beam `beamnum`'s [ngroups][nsamples][nfreq][nchans] bytes of the
// [ngroups][nbeams][nfreq][nsamples][nchans] block (transpose_to_dada.cpp:19-48).
// Both buffers on the host, or both on the device.
void do_transpose(RawBytes& transposed_data, RawBytes& input_data, std::uint32_t nchans, std::uint32_t nsamples,
                  std::uint32_t nfreq, std::uint32_t beamnum, std::uint32_t nbeams, std::uint32_t ngroups);

}  // namespace transpose

template <class HandlerType>
class TransposeToDada {
public:
    TransposeToDada(std::size_t numbeams, std::vector<std::shared_ptr<HandlerType>> handler)
        : _numbeams((std::uint32_t)numbeams), _handler(std::move(handler)), _nchans(128), _nsamples(64),
          _nfreq(32), _ngroups(10)
    {
    }
    ~TransposeToDada() {}
    TransposeToDada(TransposeToDada const&) = delete;

    void init(RawBytes& block)
    {
        for (std::uint32_t ii = 0; ii < _numbeams; ++ii) (*_handler[ii]).init(block);
    }

    bool operator()(RawBytes& block)
    {
        const std::size_t transpose_size = (std::size_t)_nchans * _nsamples * _nfreq * _ngroups;
        const std::size_t need = transpose_size * _numbeams;
        if (block.total_bytes() < need)
            throw std::runtime_error("TransposeToDada: block holds " + std::to_string(block.total_bytes()) +
                                     " bytes, the geometry needs " + std::to_string(need));
        if (_handler.size() < _numbeams) throw std::runtime_error("TransposeToDada: fewer handlers than beams");
        const void* src = block.ptr();
        if (!block.on_device()) {
            _input.resize(need);
            FBF_ERROR_CHECK(fbf_memcpy_h2d(_input.data(), block.ptr(), need, nullptr));
            src = _input.data();
        }
        _output.resize(need);
        FBF_ERROR_CHECK(fbf_tuse_transpose(src, block.on_device() ? block.total_bytes() : need, _output.data(), need,
                                           (int)_nchans, (int)_nsamples, (int)_nfreq, (int)_numbeams,
                                           (int)_ngroups, 0, (int)_numbeams, nullptr));
        _host.resize(need);
        FBF_ERROR_CHECK(fbf_memcpy_d2h(_host.data(), _output.data(), need, nullptr));
        FBF_ERROR_CHECK(fbf_stream_synchronize(nullptr));
        std::vector<std::thread> threads;
        for (std::uint32_t ii = 0; ii < _numbeams; ++ii) {
            threads.emplace_back([this, ii, transpose_size]() {
                RawBytes transpose(_host.data() + ii * transpose_size, transpose_size, 0);
                transpose.used_bytes(transpose.total_bytes());
                (*_handler[ii])(transpose);
            });
        }
        for (auto& t : threads) t.join();
        return false;
    }

    void set_nbeams(const int nbeams) { _numbeams = (std::uint32_t)nbeams; }
    void set_ngroups(const int ngroups) { _ngroups = (std::uint32_t)ngroups; }
    void set_nchans(const int nchans) { _nchans = (std::uint32_t)nchans; }
    void set_nsamples(const int nsamples) { _nsamples = (std::uint32_t)nsamples; }
    void set_nfreq(const int nfreq) { _nfreq = (std::uint32_t)nfreq; }

    std::uint32_t nchans() { return _nchans; }
    std::uint32_t nsamples() { return _nsamples; }
    std::uint32_t nfreq() { return _nfreq; }
    std::uint32_t ngroups() { return _ngroups; }
    std::uint32_t nbeams() { return _numbeams; }

private:
    std::uint32_t _numbeams;
    std::vector<std::shared_ptr<HandlerType>> _handler;
    std::uint32_t _nchans;
    std::uint32_t _nsamples;
    std::uint32_t _nfreq;
    std::uint32_t _ngroups;
    DeviceVector<char> _input;
    DeviceVector<char> _output;
    std::vector<char> _host;
};

}  // namespace tuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif
