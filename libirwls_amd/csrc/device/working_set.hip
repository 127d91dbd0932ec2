// Working-set helpers of full IRWLS training on gfx950.
//
// The working-set kernel block Kw = K(X_SW, X_SW) is materialised once per
// outer iteration; every kernel value the inner IRWLS needs (the KKT matrix,
// the inner error update, G13) is a gather from it, so the inner loop is
// memory-bound gathers instead of kernel re-evaluation.
// Reference loops replaced: src/full-train.c:169-186 (H), :236-249 (inner e),
// :313-344 (G13).
#include "common.h"
#include "working_set.h"

using namespace irwls;

namespace {

// H column-major N x N (symmetric), N = nS1 + 1, leading dimension ldh.
__global__ void build_kkt_kernel(const double *__restrict__ Kw, int ldk, const int32_t *__restrict__ S1, int nS1,
                                 const double *__restrict__ ys, const double *__restrict__ as,
                                 double *__restrict__ H, int ldh) {
    const int N = nS1 + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // row (contiguous)
    const int j = blockIdx.y;                               // column
    if (i >= N) return;
    build_kkt_at(Kw, ldk, S1, nS1, ys, as, H, ldh, i, j);
}

// One wave per output row; lanes stride the gathered columns.
__global__ __launch_bounds__(256) void gemv_gather_kernel(
    const double *__restrict__ Kw, int ldk, const int32_t *__restrict__ rows, int nr,
    const int32_t *__restrict__ cols, int nc, const double *__restrict__ w, double a, double b, double cst,
    const double *__restrict__ rsc, double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nr) return;
    const int ri = rows ? rows[i] : i;
    const double *krow = Kw + (size_t)ri * ldk;
    double acc = 0.0;
    for (int j = lane; j < nc; j += 64) acc += krow[cols ? cols[j] : j] * w[j];
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        const double sc = rsc ? rsc[i] : 1.0;
        out[i] = (a == 0.0 ? 0.0 : a * out[i]) + sc * (b * acc) + cst;
    }
}

// gemv_gather_kernel with the column count and the constant read from device
// memory (written by inner_update_kernel): out[i] = a out[i] + b (K w)_i + cst.
__global__ __launch_bounds__(256) void gemv_gather_dn_kernel(const double *__restrict__ Kw, int ldk, int nr,
                                                             const int32_t *__restrict__ cols,
                                                             const double *__restrict__ w,
                                                             const double *__restrict__ nc_cst, double a, double b,
                                                             double *__restrict__ out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nr) return;
    gemv_dn_row(Kw, ldk, cols, w, nc_cst, a, b, out, i);
}

// Right-hand side of the KKT system (src/full-train.c:160-168).
__global__ void kkt_rhs_kernel(const int32_t *__restrict__ S1, int nS1, const double *__restrict__ G13, int use_g13,
                               const double *__restrict__ gin, double last, double *__restrict__ rhs) {
    kkt_rhs_at(S1, nS1, G13, use_g13, gin, last, rhs, blockIdx.x * blockDim.x + threadIdx.x);
}

// The weights after a KKT solve and the error-update operands
// (src/full-train.c:205-249), one work-group (inner_update_wg).
__global__ __launch_bounds__(1024) void inner_update_kernel(const double *__restrict__ x,
                                                            const int32_t *__restrict__ S1, int nS1,
                                                            const int32_t *__restrict__ S3, int nS3,
                                                            const double *__restrict__ y, double C, int l,
                                                            double *__restrict__ beta, double *__restrict__ bnew,
                                                            int32_t *__restrict__ cols, double *__restrict__ w,
                                                            double *__restrict__ nc_cst) {
    inner_update_wg(x, S1, nS1, S3, nS3, y, C, l, beta, bnew, cols, w, nc_cst);
}

// y = A x (trans 0) or A^T x (trans 1) for a column-major m x n matrix A.
// trans 0: lanes over 64 rows; the columns are cut into nsl slices, one per
// wave (four per work-group along y), each summed in column order into
// part[slice * m + i]; gemv_finish adds the slices in slice order.
__global__ __launch_bounds__(256) void gemv_n_kernel(const double *__restrict__ A, int m, int n, int lda,
                                                     const double *__restrict__ x, int per,
                                                     double *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    const int slice = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int j0 = slice * per, j1 = min(n, j0 + per);
    if (i >= m) return;
    double acc = 0.0;
#pragma unroll 8
    for (int j = j0; j < j1; j++) acc = fma(A[(size_t)j * lda + i], x[j], acc);
    part[(size_t)slice * m + i] = acc;
}

__global__ void gemv_finish(const double *__restrict__ part, int m, int nsl, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int k = 0; k < nsl; k++) s += part[(size_t)k * m + i];
    y[i] = s;
}

__global__ __launch_bounds__(256) void gemv_t_kernel(const double *__restrict__ A, int m, int n, int lda,
                                                     const double *__restrict__ x, double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    const double *col = A + (size_t)j * lda;
    double acc = 0.0;
    for (int i = lane; i < m; i += 64) acc += col[i] * x[i];
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) y[j] = acc;
}

}  // namespace

extern "C" {

void irwls_dev_gemv(const double *A, int m, int n, int lda, int trans, const double *x, double *y) {
    if (m <= 0 || n <= 0) return;
    TimerScope ts(IRWLS_T_GEMV, 2.0 * m * (double)n, 8.0 * m * (double)n);
    if (!trans) {
        // at most 32 columns per wave
        const int groups = n <= 128 ? 1 : min(16, (n + 127) / 128);
        const int nsl = 4 * groups;
        const int per = (n + nsl - 1) / nsl;
        double *part = (double *)scratch(0, (size_t)nsl * m * sizeof(double));
        ts.launch(gemv_n_kernel, dim3((m + 63) / 64, groups), dim3(256), 0, A, m, n, lda, x, per, part);
        ts.launch(gemv_finish, dim3((m + 255) / 256), dim3(256), 0, part, m, nsl, y);
    } else {
        ts.launch(gemv_t_kernel, dim3((n + 3) / 4), dim3(256), 0, A, m, n, lda, x, y);
    }
}


void irwls_dev_build_kkt(const double *Kw, int ldk, const int32_t *S1, int nS1, const double *ys, const double *as,
                         double *H, int ldh) {
    const int N = nS1 + 1;
    TimerScope ts(IRWLS_T_KKT, 0.0, 8.0 * (double)N * N * 2.0);
    ts.launch(build_kkt_kernel, dim3((N + 255) / 256, N), dim3(256), 0, Kw, ldk, S1, nS1, ys, as, H, ldh);
}

void irwls_dev_kkt_rhs(const int32_t *S1, int nS1, const double *G13, int use_g13, const double *gin, double last,
                       double *rhs) {
    hipLaunchKernelGGL(kkt_rhs_kernel, dim3((nS1 + 1 + 255) / 256), dim3(256), 0, stream(), S1, nS1, G13, use_g13,
                       gin, last, rhs);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_inner_update(const double *x, const int32_t *S1, int nS1, const int32_t *S3, int nS3, const double *y,
                            double C, int l, double *beta, double *bnew, int32_t *cols, double *w, double *nc_cst) {
    hipLaunchKernelGGL(inner_update_kernel, dim3(1), dim3(1024), 0, stream(), x, S1, nS1, S3, nS3, y, C, l, beta,
                       bnew, cols, w, nc_cst);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_gemv_gather_dn(const double *Kw, int ldk, int nr, const int32_t *cols, const double *w,
                              const double *nc_cst, double a, double b, double *out) {
    if (nr <= 0) return;
    TimerScope ts(IRWLS_T_GEMV, 0.0, 0.0);
    ts.launch(gemv_gather_dn_kernel, dim3((nr + 3) / 4), dim3(256), 0, Kw, ldk, nr, cols, w, nc_cst, a, b, out);
}

void irwls_dev_gemv_gather(const double *Kw, int ldk, const int32_t *rows, int nr, const int32_t *cols, int nc,
                           const double *w, double a, double b, double cst, const double *rs, double *out) {
    if (nr <= 0) return;
    TimerScope ts(IRWLS_T_GEMV, 2.0 * nr * (double)nc, 8.0 * nr * (double)nc);
    ts.launch(gemv_gather_kernel, dim3((nr + 3) / 4), dim3(256), 0, Kw, ldk, rows, nr, cols, nc, w, a, b, cst, rs,
              out);
}

}  // extern "C"
