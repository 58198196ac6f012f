// sg_error.hpp -- thread-local last-error text behind the sg_status codes.
#pragma once

#include <cstdio>
#include <string>

namespace sg {

std::string &last_error();

inline int fail(int code, const std::string &msg)
{
    last_error() = msg;
    return code;
}

}  // namespace sg

// Evaluates a HIP call; on failure records where and returns SG_ERR_HIP.
#define SG_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            char b_[512];                                                                 \
            snprintf(b_, sizeof b_, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                                 \
            return sg::fail(SG_ERR_HIP, b_);                                              \
        }                                                                                 \
    } while (0)
