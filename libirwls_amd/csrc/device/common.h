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

// Lane `lane`'s value of v, wave-uniform (v_readlane on both halves).
__device__ __forceinline__ double rdl(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The library stream: every kernel and copy of the library is issued here.
hipStream_t stream();

// Timing brackets around launches of one kind (no-ops unless enabled).
struct TimerScope {
    int kind;
    double flops, bytes;
    bool on;
    // any_thread: also on worker threads (for a launch that only one thread
    // at a time can make, such as a batched solve)
    TimerScope(int kind, double flops, double bytes, bool any_thread = false);
    ~TimerScope();
};

// Scratch device buffer owned by the context, grown on demand.
void *scratch(int slot, size_t bytes);

}  // namespace irwls
