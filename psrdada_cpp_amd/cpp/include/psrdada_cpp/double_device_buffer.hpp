// DoubleDeviceBuffer<T> (double_device_buffer.cuh:9-10) over DeviceVector<T>.
#ifndef PSRDADA_CPP_DOUBLE_DEVICE_BUFFER_HPP
#define PSRDADA_CPP_DOUBLE_DEVICE_BUFFER_HPP

#include "psrdada_cpp/device_vector.hpp"
#include "psrdada_cpp/double_buffer.hpp"

namespace psrdada_cpp {

template <typename T>
using DoubleDeviceBuffer = DoubleBuffer<DeviceVector<T>>;

}  // namespace psrdada_cpp

#endif
