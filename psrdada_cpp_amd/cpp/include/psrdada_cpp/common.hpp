// Common types of the C++ operator layer.
//
// The operator classes speak plain C++ over the C ABI of libfbfuse_hip
// (include/fbfuse_hip.h); no HIP header is needed to use them.  Complex int8
// samples and delay models use the same two-member layout as the reference's
// CUDA vector types (char2 / float2), declared here in namespace psrdada_cpp.
#ifndef PSRDADA_CPP_COMMON_HPP
#define PSRDADA_CPP_COMMON_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "fbfuse_hip.h"

namespace psrdada_cpp {

struct char2 {
    int8_t x;  // real
    int8_t y;  // imaginary
};
static_assert(sizeof(char2) == 2, "char2 must be two bytes");

struct float2 {
    float x;
    float y;
};
static_assert(sizeof(float2) == 8, "float2 must be eight bytes");

// A hipStream_t converts implicitly to this handle.
typedef fbf_stream_t StreamType;

// Turns a C ABI status into a std::runtime_error (cuda_utils.hpp:25-47 does
// the same for cudaError_t).
inline void fbf_check(int status, const char* what)
{
    if (status != FBF_OK) {
        const char* msg = fbf_last_error();
        throw std::runtime_error(std::string(what) + " failed with status " + std::to_string(status) + ": " +
                                 (msg ? msg : ""));
    }
}

}  // namespace psrdada_cpp

#define FBF_ERROR_CHECK(expr) ::psrdada_cpp::fbf_check((expr), #expr)

#endif
