// Device bodies of the working-set kernels of full IRWLS training, shared by
// the single-problem kernels (working_set.hip) and the batched round
// (cholesky.hip), so both compute every value with the same operations.
// Reference loops: src/full-train.c:160-186 (right-hand side and H),
// :205-249 (weights after a solve, inner error update).
#pragma once

#include "common.h"

namespace irwls {

// H[j*ldh + i] of the KKT matrix (N = nS1 + 1)
__device__ __forceinline__ void build_kkt_at(const double *__restrict__ Kw, int ldk, const int32_t *__restrict__ S1,
                                             int nS1, const double *__restrict__ ys, const double *__restrict__ as,
                                             double *__restrict__ H, int ldh, int i, int j) {
    double v;
    if (i < nS1 && j < nS1) {
        v = Kw[(size_t)S1[i] * ldk + S1[j]] * ys[i] * ys[j];
        if (i == j) v += 1.0 / as[i];
    } else if (i < nS1) {
        v = ys[i];          // last column: y border
    } else if (j < nS1) {
        v = ys[j];          // last row: y border
    } else {
        v = 0.0;
    }
    H[(size_t)j * ldh + i] = v;
}

// rhs_i = 1 - G13_i - GIN[S1_i] (i < nS1; G13 read as zero when use_g13 is 0), rhs_nS1 = last
__device__ __forceinline__ void kkt_rhs_at(const int32_t *__restrict__ S1, int nS1, const double *__restrict__ G13,
                                           int use_g13, const double *__restrict__ gin, double last,
                                           double *__restrict__ rhs, int i) {
    if (i < nS1) rhs[i] = 1.0 - (use_g13 ? G13[i] : 0.0) - gin[S1[i]];
    else if (i == nS1) rhs[i] = last;
}

// One work-group: bnew = 0, bnew[S1_i] = x_i y[S1_i], bnew[S3_o] = C y[S3_o],
// bnew_l = x_nS1; then the columns j < l whose weight changed, in ascending
// order, with w = bnew_j - beta_j, their count, the constant
// -(bnew_l - beta_l), and beta = bnew.  nc_cst: count (int32 at word 0)
// and constant (word 1).
__device__ __forceinline__ void inner_update_wg(const double *__restrict__ x, const int32_t *__restrict__ S1,
                                                int nS1, const int32_t *__restrict__ S3, int nS3,
                                                const double *__restrict__ y, double C, int l,
                                                double *__restrict__ beta, double *__restrict__ bnew,
                                                int32_t *__restrict__ cols, double *__restrict__ w,
                                                double *__restrict__ nc_cst) {
    __shared__ int part[16];
    __shared__ int base_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
    for (int j = t; j <= l; j += blockDim.x) bnew[j] = 0.0;
    __syncthreads();
    for (int i = t; i < nS1; i += blockDim.x) bnew[S1[i]] = x[i] * y[S1[i]];
    for (int o = t; o < nS3; o += blockDim.x) bnew[S3[o]] = C * y[S3[o]];
    if (t == 0) {
        bnew[l] = x[nS1];
        base_s = 0;
    }
    __syncthreads();
    for (int j0 = 0; j0 < l; j0 += blockDim.x) {
        const int j = j0 + t;
        const bool ch = (j < l) && (bnew[j] != beta[j]);
        const unsigned long long m = __ballot(ch);
        const int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) part[wv] = __popcll(m);
        __syncthreads();
        int off = base_s;
        for (int k = 0; k < wv; k++) off += part[k];
        if (ch) {
            cols[off + below] = j;
            w[off + below] = bnew[j] - beta[j];
        }
        __syncthreads();
        if (t == 0) {
            int s = 0;
            for (int k = 0; k < nw; k++) s += part[k];
            base_s += s;
        }
        __syncthreads();
    }
    if (t == 0) {
        *reinterpret_cast<int32_t *>(nc_cst) = base_s;
        nc_cst[1] = -(bnew[l] - beta[l]);
    }
    __syncthreads();
    for (int j = t; j <= l; j += blockDim.x) beta[j] = bnew[j];
}

// Row i of out = a out + b (K w) + cst over the columns cols[0..nc), nc and
// cst read from nc_cst; one wave, lanes stride the columns.
__device__ __forceinline__ void gemv_dn_row(const double *__restrict__ Kw, int ldk, const int32_t *__restrict__ cols,
                                            const double *__restrict__ w, const double *__restrict__ nc_cst,
                                            double a, double b, double *__restrict__ out, int i) {
    const int lane = threadIdx.x & 63;
    const int nc = *reinterpret_cast<const int32_t *>(nc_cst);
    const double *krow = Kw + (size_t)i * ldk;
    double acc = 0.0;
    for (int j = lane; j < nc; j += 64) acc += krow[cols[j]] * w[j];
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) out[i] = (a == 0.0 ? 0.0 : a * out[i]) + (b * acc) + nc_cst[1];
}

}  // namespace irwls
