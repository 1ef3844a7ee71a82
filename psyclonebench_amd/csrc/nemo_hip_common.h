// Shared host-side helpers for the HIP library: error checking with the
// reference's print-and-exit convention (psykal_c_opencl/opencl_utils.c:
// check_status prints the failing call and calls exit(1)).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define NEMO_HIP_CHECK(call)                                                   \
    do {                                                                       \
        hipError_t nemo_err_ = (call);                                         \
        if (nemo_err_ != hipSuccess) {                                         \
            std::fprintf(stderr, "nemolite2d_hip: %s failed at %s:%d: %s\n",   \
                         #call, __FILE__, __LINE__,                            \
                         hipGetErrorString(nemo_err_));                        \
            std::exit(1);                                                      \
        }                                                                      \
    } while (0)

#define NEMO_FATAL(...)                                                        \
    do {                                                                       \
        std::fprintf(stderr, "nemolite2d_hip: " __VA_ARGS__);                  \
        std::fprintf(stderr, "\n");                                            \
        std::exit(1);                                                          \
    } while (0)

#define NEMO_LAUNCH_CHECK() NEMO_HIP_CHECK(hipGetLastError())
