// Kernel-block x vector product that allocates nothing
// (irwls_dev_kernel_matvec_noalloc).
//
// irwls_dev_kernel_matvec keeps the column splits' partial sums in a scratch
// slot that it grows with hipMalloc.  Callers that cannot allocate (work
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) get the same kernels with the partial sums in a fixed device
// array of this file instead.  Each GPU holds its own copy of the array from
// the first use until the process ends; host code reaches it only through
// the per-device address from hipGetSymbolAddress, never through its name.
#include "common.h"

#include <mutex>

// Partial sums: at most IRWLS_NOALLOC_SCRATCH_BYTES (16 MiB), enough for
// every RBF shape on a 256-CU part (nsplit x rows <= 2^21 doubles) and for
// the linear kernel up to d = 32768 features.
static __device__ __align__(256) double scratch[IRWLS_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];   // device address of `scratch`, per device id

// The current device's address of the array, looked up once per device.
void *device_scratch() {
    int dev = 0;
    IRWLS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        fprintf(stderr, "libirwls: device id %d out of range for the no-alloc scratch table\n", dev);
        exit(2);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "libirwls: hipGetSymbolAddress of the no-alloc scratch failed: %s\n",
                    hipGetErrorString(e));
            exit(2);
        }
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

}  // namespace

namespace irwls {

void noalloc_setup() { (void)device_scratch(); }

}  // namespace irwls

extern "C" {

int irwls_dev_kernel_matvec_noalloc(const irwls_rows *r, const irwls_rows *c, int d, int kernel, double gamma,
                                    const double *w, double a, double b, double cst, const double *rs,
                                    double *out) {
    if (!irwls::dev_initialised()) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matvec_noalloc needs irwls_dev_init first\n");
        return -1;
    }
    if (!r || !c || r->n < 0 || c->n < 0 || d < 0) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matvec_noalloc: bad arguments\n");
        return -1;
    }
    if (r->n == 0) return 0;
    if (c->n == 0) {
        irwls::kmv_empty_cols(r->n, a, cst, out);
        return 0;
    }
    const bool linear = kernel == 0 && d > 0;
    if (!c->X || (!linear && !r->X)) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matvec_noalloc takes dense columns (and dense rows for the "
                        "RBF kernel)\n");
        return -1;
    }
    const size_t need = irwls::kmv_part_bytes(r, c, d, kernel);
    if (need > (size_t)IRWLS_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matvec_noalloc: %d x %d (d = %d) needs %zu bytes of partial "
                        "sums, more than its fixed %d\n", r->n, c->n, d, need, IRWLS_NOALLOC_SCRATCH_BYTES);
        return -1;
    }
    // every partial sum the launch reads is written by it first: nothing to clear
    irwls::kmv_dense_cols_run(r, c, d, kernel, gamma, w, a, b, cst, rs, out,
                              need ? (double *)device_scratch() : nullptr, false);
    return 0;
}

}  // extern "C"
