// Shared definitions of the gfx950 device layer.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
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

// Event pair `which` (0 start, 1 stop) of timer slot idx.
hipEvent_t timer_event(int idx, int which);

// Timing of launches of one kind (no-ops unless enabled; a negative kind is
// never timed).  A scope's kernels are launched through launch(): the first
// kernel's dispatch carries the start event and every kernel's the stop
// event (hipExtLaunchKernel), so the time is the kernels' own, from the
// first one's start to the last one's end, as a kernel trace reports it.
// With markers = true the scope instead records an event before and after
// whatever it encloses (copies, graph launches, calls of other functions).
struct TimerScope {
    int kind;
    double flops, bytes;
    bool on, markers;
    int idx;   // event pair reserved for this scope
    int nk;    // kernels launched through launch()
    // any_thread: also on worker threads (for a launch that only one thread
    // at a time can make, such as a batched solve)
    TimerScope(int kind, double flops, double bytes, bool any_thread = false, bool markers = false);
    ~TimerScope();
    template <typename F, typename... A>
    void launch(F kernel, dim3 grid, dim3 block, unsigned shm, A... args) {
        if (on && !markers) {
            hipExtLaunchKernelGGL(kernel, grid, block, shm, stream(), nk == 0 ? timer_event(idx, 0) : nullptr,
                                  timer_event(idx, 1), 0u, args...);
            nk++;
        } else {
            hipLaunchKernelGGL(kernel, grid, block, shm, stream(), args...);
        }
        IRWLS_LAUNCH_CHECK();
    }
};

// Scratch device buffer owned by the context, grown on demand.
void *scratch(int slot, size_t bytes);

// Kernel-block x vector product over dense columns (c->X set, r->n and
// c->n > 0) with a caller-given partial-sum buffer of kmv_part_bytes bytes
// (0: none needed); timed = false records no timer events.
size_t kmv_dense_part_bytes(int rn, int cn);
size_t kmv_primal_part_bytes(int cn, int d);
inline size_t kmv_part_bytes(const irwls_rows *r, const irwls_rows *c, int d, int kernel) {
    return (kernel == 0 && d > 0) ? kmv_primal_part_bytes(c->n, d) : kmv_dense_part_bytes(r->n, c->n);
}
void kmv_dense_cols_run(const irwls_rows *r, const irwls_rows *c, int d, int kernel, double gamma, const double *w,
                        double a, double b, double cst, const double *rs, double *out, double *part, bool timed);
void kmv_empty_cols(int n, double a, double cst, double *out);
// True once irwls_dev_init has run.
bool dev_initialised();
// Looks up the current device's address of the fixed partial-sum array of
// irwls_dev_kernel_matvec_noalloc (called by irwls_dev_init).
void noalloc_setup();

// Small device-to-host read that does not drain the stream: the copy is
// queued into a pinned buffer of the calling thread (at most kReadbackMax
// bytes) and marked by an event; readback_wait blocks until it has landed
// and returns the buffer.  One read in flight per thread.
constexpr size_t kReadbackMax = (size_t)64 << 10;
void readback_start(const void *src, size_t bytes);
const void *readback_wait();

// Compute units of the current device.
int device_cus();

// Held around a stream capture: no timer event is recorded or read meanwhile
// (an event last recorded on a capturing stream cannot be used).
void capture_lock();
void capture_unlock();

// Column splits of a kernel-block x vector launch over rt row tiles and ct
// column tiles, for a kernel of which `per_cu` work-groups fit on a compute
// unit: the fewest splits (powers of two) that give each resident
// work-group at least 16 row-tile x split items over the launch, while each
// split keeps at least two column tiles.  Depends on the shape only, so a
// given shape always sums in the same order.
inline void column_splits(int rt, int ct, int per_cu, int *tps, int *nsplit) {
    const long slots = (long)device_cus() * per_cu;
    int ns = 1;
    while ((long)rt * ns < 16 * slots && ct >= 4 * ns) ns *= 2;
    *tps = (ct + ns - 1) / ns;
    *nsplit = (ct + *tps - 1) / *tps;
}

}  // namespace irwls
