// Non-owning view of a host or device buffer (raw_bytes.hpp:19-69).
#ifndef PSRDADA_CPP_RAW_BYTES_HPP
#define PSRDADA_CPP_RAW_BYTES_HPP

#include <cstddef>

namespace psrdada_cpp {

class RawBytes {
public:
    RawBytes(char* ptr, std::size_t total, std::size_t used = 0, bool device_memory = false)
        : _ptr(ptr), _total_bytes(total), _used_bytes(used), _on_device(device_memory)
    {
    }
    RawBytes(RawBytes const&) = delete;
    ~RawBytes() = default;

    std::size_t total_bytes() const { return _total_bytes; }
    std::size_t used_bytes() const { return _used_bytes; }
    void used_bytes(std::size_t used) { _used_bytes = used; }
    char* ptr() { return _ptr; }
    bool on_device() const { return _on_device; }

private:
    char* _ptr;
    std::size_t _total_bytes;
    std::size_t _used_bytes;
    bool _on_device;
};

}  // namespace psrdada_cpp

#endif
