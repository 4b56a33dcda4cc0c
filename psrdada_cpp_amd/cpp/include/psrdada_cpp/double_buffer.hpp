// a/b double buffer over two vectors (double_buffer.cuh:20-79,
// detail/double_buffer.cu:18-66): swap() exchanges the contents of a and b.
#ifndef PSRDADA_CPP_DOUBLE_BUFFER_HPP
#define PSRDADA_CPP_DOUBLE_BUFFER_HPP

#include <cstddef>

namespace psrdada_cpp {

template <typename VectorType>
class DoubleBuffer {
public:
    DoubleBuffer() = default;
    ~DoubleBuffer() = default;
    DoubleBuffer(DoubleBuffer const&) = delete;

    void resize(std::size_t size)
    {
        _buf0.resize(size);
        _buf1.resize(size);
    }
    void resize(std::size_t size, typename VectorType::value_type fill_value)
    {
        _buf0.resize(size, fill_value);
        _buf1.resize(size, fill_value);
    }
    std::size_t size() const { return _buf0.size(); }
    void swap() { _buf0.swap(_buf1); }
    VectorType& a() { return _buf0; }
    VectorType& b() { return _buf1; }
    typename VectorType::value_type* a_ptr() { return _buf0.data(); }
    typename VectorType::value_type* b_ptr() { return _buf1.data(); }

private:
    VectorType _buf0;
    VectorType _buf1;
};

}  // namespace psrdada_cpp

#endif
