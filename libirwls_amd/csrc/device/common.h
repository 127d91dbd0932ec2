// Shared definitions of the gfx950 device layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>

#include "irwls_device.h"

#define IRWLS_HIP(call)                                                        \
    do {                                                                       \
        hipError_t err_ = (call);                                              \
        if (err_ != hipSuccess) {                                              \
            fprintf(stderr, "libirwls: HIP error %s at %s:%d: %s\n",          \
                    hipGetErrorName(err_), __FILE__, __LINE__,                 \
                    hipGetErrorString(err_));                                  \
            exit(2);                                                           \
        }                                                                      \
    } while (0)

#define IRWLS_LAUNCH_CHECK() IRWLS_HIP(hipGetLastError())

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace irwls {

// The library stream: every kernel and copy of the library is issued here.
hipStream_t stream();

// Timing brackets around launches of one kind (no-ops unless enabled).
struct TimerScope {
    int kind;
    double flops, bytes;
    bool on;
    TimerScope(int kind, double flops, double bytes);
    ~TimerScope();
};

// Scratch device buffer owned by the context, grown on demand.
void *scratch(int slot, size_t bytes);

}  // namespace irwls
