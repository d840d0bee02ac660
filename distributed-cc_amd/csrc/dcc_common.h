// Internal helpers shared by the libdcc translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "dcc.h"

namespace dcc {

void set_error(const std::string &msg);
void clear_error();

inline hipStream_t as_stream(dcc_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace dcc

// Return a nonzero status from an entry point when a HIP call fails.
#define DCC_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t err_ = (call);                                                          \
        if (err_ != hipSuccess) {                                                          \
            dcc::set_error(std::string(#call) + ": " + hipGetErrorString(err_));            \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

#define DCC_CHECK(cond, msg)                                                               \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            dcc::set_error(std::string(__func__) + ": " + (msg));                           \
            return 2;                                                                      \
        }                                                                                  \
    } while (0)

// Launch status: a launch error is returned as status 1.
#define DCC_LAUNCH_CHECK()                                                                 \
    do {                                                                                   \
        hipError_t err_ = hipGetLastError();                                               \
        if (err_ != hipSuccess) {                                                          \
            dcc::set_error(std::string(__func__) + " launch: " + hipGetErrorString(err_));  \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)
