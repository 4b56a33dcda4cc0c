// Owning device array, the storage behind DoubleDeviceBuffer.
//
// Replaces the thrust::device_vector the reference uses
// (double_device_buffer.cuh:9-10): resize keeps the existing elements, new
// elements are value-initialised (or set to the fill value), copies are deep,
// swap exchanges storage.  Memory comes from fbf_malloc (hipMalloc).
#ifndef PSRDADA_CPP_DEVICE_VECTOR_HPP
#define PSRDADA_CPP_DEVICE_VECTOR_HPP

#include <algorithm>
#include <utility>
#include <vector>

#include "psrdada_cpp/common.hpp"

namespace psrdada_cpp {

template <typename T>
class DeviceVector {
public:
    typedef T value_type;

    DeviceVector() = default;
    explicit DeviceVector(std::size_t n) { resize(n); }
    DeviceVector(std::size_t n, T const& fill) { resize(n, fill); }
    explicit DeviceVector(std::vector<T> const& host) { assign(host.data(), host.size()); }
    DeviceVector(DeviceVector const& other) { *this = other; }
    DeviceVector(DeviceVector&& other) noexcept { swap(other); }
    ~DeviceVector()
    {
        if (_ptr) fbf_free(_ptr);
    }

    DeviceVector& operator=(DeviceVector const& other)
    {
        if (this != &other) {
            reserve_exact(other._size, false);
            _size = other._size;
            if (_size) FBF_ERROR_CHECK(fbf_memcpy_d2d(_ptr, other._ptr, bytes(), nullptr));
            FBF_ERROR_CHECK(fbf_device_synchronize());
        }
        return *this;
    }
    DeviceVector& operator=(DeviceVector&& other) noexcept
    {
        swap(other);
        return *this;
    }
    DeviceVector& operator=(std::vector<T> const& host)
    {
        assign(host.data(), host.size());
        return *this;
    }

    std::size_t size() const { return _size; }
    bool empty() const { return _size == 0; }
    std::size_t bytes() const { return _size * sizeof(T); }
    T* data() { return _ptr; }
    T const* data() const { return _ptr; }

    void resize(std::size_t n) { resize_impl(n, nullptr); }
    void resize(std::size_t n, T const& fill) { resize_impl(n, &fill); }

    void swap(DeviceVector& other) noexcept
    {
        std::swap(_ptr, other._ptr);
        std::swap(_size, other._size);
        std::swap(_capacity, other._capacity);
    }

    // Host <-> device copies (synchronous on the given stream).
    void assign(T const* host, std::size_t n, StreamType stream = nullptr)
    {
        reserve_exact(n, false);
        _size = n;
        if (n) FBF_ERROR_CHECK(fbf_memcpy_h2d(_ptr, host, bytes(), stream));
        FBF_ERROR_CHECK(stream ? fbf_stream_synchronize(stream) : fbf_device_synchronize());
    }
    std::vector<T> to_host(StreamType stream = nullptr) const
    {
        std::vector<T> out(_size);
        if (_size) FBF_ERROR_CHECK(fbf_memcpy_d2h(out.data(), _ptr, bytes(), stream));
        FBF_ERROR_CHECK(stream ? fbf_stream_synchronize(stream) : fbf_device_synchronize());
        return out;
    }

private:
    void reserve_exact(std::size_t n, bool keep)
    {
        if (n <= _capacity) return;
        void* p = nullptr;
        FBF_ERROR_CHECK(fbf_malloc(&p, n * sizeof(T)));
        if (keep && _size) {
            FBF_ERROR_CHECK(fbf_memcpy_d2d(p, _ptr, bytes(), nullptr));
            FBF_ERROR_CHECK(fbf_device_synchronize());
        }
        if (_ptr) fbf_free(_ptr);
        _ptr = static_cast<T*>(p);
        _capacity = n;
    }

    void resize_impl(std::size_t n, T const* fill)
    {
        const std::size_t old = _size;
        reserve_exact(n, true);
        _size = n;
        if (n > old) {
            const std::size_t extra = n - old;
            if (fill) {
                std::vector<T> host(extra, *fill);
                FBF_ERROR_CHECK(fbf_memcpy_h2d(_ptr + old, host.data(), extra * sizeof(T), nullptr));
            } else {
                FBF_ERROR_CHECK(fbf_memset(_ptr + old, 0, extra * sizeof(T), nullptr));
            }
            FBF_ERROR_CHECK(fbf_device_synchronize());
        }
    }

    T* _ptr = nullptr;
    std::size_t _size = 0;
    std::size_t _capacity = 0;
};

}  // namespace psrdada_cpp

#endif
