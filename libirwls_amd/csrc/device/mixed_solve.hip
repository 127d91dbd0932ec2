// Mixed-precision SPD solve for two right-hand sides: preconditioned
// conjugate gradients in FP64 on the FP64 matrix, with the FP32 Cholesky
// factor (FP64 diagonal tiles) as preconditioner M^-1 = Z Z^T, Z = L^-T.
//
// The FP64 iteration recovers FP64 residuals from the FP32 factor; CG
// converges for any SPD preconditioner, so a factor of an earlier matrix
// (same index set, other diagonal) still serves, at the cost of more
// iterations.  Every matrix-vector product is one wave per output entry over
// a contiguous column: A is symmetric (column i = row i), Z^T r reads the
// columns of Z from the factor's identity rows and Z W the columns of Z^T.
// Dot products leave one partial per work-group; the next kernel reduces
// them (each work-group redundantly), so no launch is spent on reductions.
// Vectors with two columns are stored [2][n].
#include "common.h"

using namespace irwls;

namespace {

// out_c(i) = sum_{k in range(i)} M[i*ld + off + k] v_c(k), c = 0, 1, with
// range [0, n) (mode 0), [0, i] (mode 1) or [i, n) (mode 2).  One
// work-group per i: its four waves take contiguous quarters of the range
// (loads unrolled four deep), added in wave order.  When dot_with is set,
// dot_with_c(i) out_c(i) goes to part[2*i + c].
template <typename T, int MODE>
__global__ __launch_bounds__(256) void coldot2(const T *__restrict__ M, int ld, size_t off, int n,
                                               const double *__restrict__ v, double *__restrict__ out,
                                               const double *__restrict__ dot_with, double *__restrict__ part) {
    __shared__ double wp[4][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x;
    const int lo = (MODE == 2) ? i : 0;
    const int hi = (MODE == 1) ? i + 1 : n;
    const int len = hi - lo;
    const int qa = lo + (int)(((long long)len * wave) / 4), qb = lo + (int)(((long long)len * (wave + 1)) / 4);
    const T *col = M + (size_t)i * ld + off;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
    for (int k = qa + lane; k < qb; k += 64) {
        const double m = (double)col[k];
        s0 = fma(m, v[k], s0);
        s1 = fma(m, v[n + k], s1);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
    }
    if (lane == 0) {
        wp[wave][0] = s0;
        wp[wave][1] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int c = threadIdx.x;
        const double t = (wp[0][c] + wp[1][c]) + (wp[2][c] + wp[3][c]);
        out[(size_t)c * n + i] = t;
        if (dot_with) part[2 * i + c] = dot_with[(size_t)c * n + i] * t;
    }
}

// Sum of np partial pairs (part[2p + c]) into red[c], by all 256 threads.
__device__ __forceinline__ void reduce_parts(const double *__restrict__ part, int np, double *red) {
    __shared__ double sh[2][256];
    double a0 = 0.0, a1 = 0.0;
    for (int p = threadIdx.x; p < np; p += 256) {
        a0 += part[2 * p];
        a1 += part[2 * p + 1];
    }
    sh[0][threadIdx.x] = a0;
    sh[1][threadIdx.x] = a1;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    red[0] = sh[0][0];
    red[1] = sh[1][0];
    __syncthreads();
}

// alpha_c = rz_c / (p_c . q_c);  x_c += alpha_c p_c;  r_c -= alpha_c q_c;
// work-group partials of |r_c|^2 to rr_part.  A column whose rz is zero
// (solved, or a zero right-hand side) is left as it is.
__global__ __launch_bounds__(256) void update_xr(int n, const double *__restrict__ pq_part, int npq,
                                                 const double *__restrict__ rz, const double *__restrict__ p,
                                                 const double *__restrict__ q, double *__restrict__ x,
                                                 double *__restrict__ r, double *__restrict__ rr_part) {
    __shared__ double wsum[2][4];
    double pq[2];
    reduce_parts(pq_part, npq, pq);
    const double al0 = (rz[0] != 0.0 && pq[0] != 0.0) ? rz[0] / pq[0] : 0.0;
    const double al1 = (rz[1] != 0.0 && pq[1] != 0.0) ? rz[1] / pq[1] : 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double r0 = 0.0, r1 = 0.0;
    if (i < n) {
        x[i] = fma(al0, p[i], x[i]);
        x[n + i] = fma(al1, p[n + i], x[n + i]);
        r0 = fma(-al0, q[i], r[i]);
        r1 = fma(-al1, q[n + i], r[n + i]);
        r[i] = r0;
        r[n + i] = r1;
    }
    double s0 = r0 * r0, s1 = r1 * r1;
    for (int o = 32; o >= 1; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wsum[0][wave] = s0;
        wsum[1][wave] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int c = threadIdx.x;
        rr_part[2 * blockIdx.x + c] = (wsum[c][0] + wsum[c][1]) + (wsum[c][2] + wsum[c][3]);
    }
}

// rz_new_c = sum of partials; beta_c = rz_new_c / rz_old_c (0 on the first
// call); p_c = z_c + beta_c p_c; work-group 0 stores rz_new in rz_out (a
// different slot from rz_old, so no work-group reads what another writes).
__global__ __launch_bounds__(256) void update_p(int n, const double *__restrict__ rz_part, int nrz,
                                                const double *__restrict__ rz_old, int first,
                                                double *__restrict__ rz_out, const double *__restrict__ z,
                                                double *__restrict__ p) {
    double rzn[2];
    reduce_parts(rz_part, nrz, rzn);
    const double b0 = (first || rz_old[0] == 0.0) ? 0.0 : rzn[0] / rz_old[0];
    const double b1 = (first || rz_old[1] == 0.0) ? 0.0 : rzn[1] / rz_old[1];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        p[i] = fma(b0, p[i], z[i]);
        p[n + i] = fma(b1, p[n + i], z[n + i]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rz_out[0] = rzn[0];
        rz_out[1] = rzn[1];
    }
}

// Work-group partials of |b_c|^2 (c = 0, 1) to part[2 * blockIdx.x + c],
// summed like the residual partials of update_xr.
__global__ __launch_bounds__(256) void sumsq2_parts(int n, const double *__restrict__ b, double *__restrict__ part) {
    __shared__ double wsum[2][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    if (i < n) {
        s0 = b[i] * b[i];
        s1 = b[n + i] * b[n + i];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wsum[0][wave] = s0;
        wsum[1][wave] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int c = threadIdx.x;
        part[2 * blockIdx.x + c] = (wsum[c][0] + wsum[c][1]) + (wsum[c][2] + wsum[c][3]);
    }
}

// Preconditioner of a system whose index set differs from the factored
// one.  inv[f] = row of the system at factor position f (or -1), map[i] =
// factor position of row i (or -1).  r is scattered into the factor's index
// space (zeros where the system has no row), Z Z^T applied there, and the
// result gathered back; rows the factor lacks take the Jacobi step r_i/A_ii.
// For rows removed from a factored SPD matrix the operator (A_old^-1)_CC is
// SPD and differs from (A_CC)^-1 by a term of rank |removed|.
__global__ void scatter_to_factor(int n, int nf, const int32_t *__restrict__ inv, const double *__restrict__ r,
                                  double *__restrict__ rf) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int i = inv[f];
    rf[f] = (i >= 0) ? r[i] : 0.0;
    rf[nf + f] = (i >= 0) ? r[n + i] : 0.0;
}

// With rs > 0 rows removed from the factored set (factor positions R), the
// kept rows' operator becomes the exact inverse of the kept block of the
// factored matrix: B_KK - B_KR S^-1 B_RK with B = Z Z^T and S = B_RR, since
// for SPD A the inverse of a principal block A_KK equals
// (A^-1)_KK - (A^-1)_KR ((A^-1)_RR)^-1 (A^-1)_RK.  U = B(:, R) (nf x rs,
// column-major) and S^-1 (rs x rs, dense) come from
// irwls_dev_pcg_schur_setup; per application y = S^-1 zf(R) (schur_apply)
// and zf -= U y over the factor's index space (schur_correct), before the
// kept rows are gathered.
constexpr int SCHUR_MAXR = 256;
__global__ __launch_bounds__(256) void schur_apply(const double *__restrict__ zf, int nf, const int32_t *__restrict__ R,
                                                   int rs, const double *__restrict__ Sinv, double *__restrict__ y) {
    // one wave per row a of S^-1 (symmetric: its row a is its column a)
    const int lane = threadIdx.x & 63, a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= rs) return;
    const double *col = Sinv + (size_t)a * rs;
    double s0 = 0.0, s1 = 0.0;
    for (int b = lane; b < rs; b += 64) {
        const int f = R[b];
        s0 = fma(col[b], zf[f], s0);
        s1 = fma(col[b], zf[nf + f], s1);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
    }
    if (lane == 0) {
        y[a] = s0;
        y[SCHUR_MAXR + a] = s1;
    }
}

// zf_c(f) -= sum_j U(f, j) y_c(j): work-group of 64 factor positions (one per
// lane), its four waves summing quarters of j, added in wave order.
__global__ __launch_bounds__(256) void schur_correct(double *__restrict__ zf, int nf, const double *__restrict__ U,
                                                     int rs, const double *__restrict__ y) {
    __shared__ double ws[2][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int f = blockIdx.x * 64 + lane;
    const int q = (rs + 3) / 4, j0 = wave * q, j1 = min(rs, j0 + q);
    double s0 = 0.0, s1 = 0.0;
    if (f < nf) {
#pragma unroll 8
        for (int j = j0; j < j1; j++) {
            const double u = U[(size_t)j * nf + f];
            s0 = fma(u, y[j], s0);
            s1 = fma(u, y[SCHUR_MAXR + j], s1);
        }
    }
    ws[0][wave][lane] = s0;
    ws[1][wave][lane] = s1;
    __syncthreads();
    if (wave < 2 && f < nf) {
        const int c = wave;
        const double t = (ws[c][0][lane] + ws[c][1][lane]) + (ws[c][2][lane] + ws[c][3][lane]);
        zf[(size_t)c * nf + f] -= t;
    }
}

// Partial sums of U(i, j) = sum_k Z(i, k) Z(R_j, k) (Z(i, k) = ZT[i*ldz + k],
// zero for k < i): work-group (bx, s, cb) takes rows 64 bx .. 64 bx + 63,
// columns 64 cb .. 64 cb + 63 and k in [SCHUR_SK s, SCHUR_SK (s+1)) from
// k = 64 bx on (a split entirely below that has no work and writes nothing);
// each thread a 4 x 4 block of (row, column) pairs, operands staged in LDS
// as FP64 (the FP32 products are exact there), partials to
// part[(s rs + j) nf + i].
constexpr int SCHUR_SK = 128, SCHUR_KC = 32;
__global__ __launch_bounds__(256) void schur_u_part(const float *__restrict__ ZT, int ldz, int nf,
                                                    const int32_t *__restrict__ R, int rs, double *__restrict__ part) {
    __shared__ double Am[SCHUR_KC][65];   // Am[kk][row] = Z(i0 + row, k0 + kk)
    __shared__ double Bm[SCHUR_KC][65];   // Bm[kk][j] = Z(R_j, k0 + kk)
    const int i0 = blockIdx.x * 64, sp = blockIdx.y, j0 = blockIdx.z * 64;
    const int kbeg = max(sp * SCHUR_SK, i0), kend = min(sp * SCHUR_SK + SCHUR_SK, nf);
    if (kbeg >= kend) return;
    const int t = threadIdx.x, tr = t & 15, tc = t >> 4;   // rows tr + 16 q, columns tc + 16 p
    double acc[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) acc[q][p] = 0.0;
    for (int k0 = kbeg; k0 < kend; k0 += SCHUR_KC) {
        __syncthreads();
        for (int e = t; e < SCHUR_KC * 64; e += 256) {
            const int kk = e & (SCHUR_KC - 1), row = e / SCHUR_KC, k = k0 + kk;
            const int i = i0 + row;
            Am[kk][row] = (i < nf && k < kend) ? (double)ZT[(size_t)i * ldz + k] : 0.0;
            Bm[kk][row] = (j0 + row < rs && k < kend) ? (double)ZT[(size_t)R[j0 + row] * ldz + k] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < SCHUR_KC; kk++) {
            double a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; q++) a[q] = Am[kk][tr + 16 * q];
#pragma unroll
            for (int p = 0; p < 4; p++) b[p] = Bm[kk][tc + 16 * p];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int p = 0; p < 4; p++) acc[q][p] = fma(a[q], b[p], acc[q][p]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = i0 + tr + 16 * q;
        if (i >= nf) continue;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int j = j0 + tc + 16 * p;
            if (j < rs) part[((size_t)sp * rs + j) * nf + i] = acc[q][p];
        }
    }
}

// U(:, j) = sum of the splits' partials (fixed order; splits entirely below
// the row's 64-row block hold nothing); S tile (64 x 64, column-major) =
// U(R, :) padded with the identity.
__global__ void schur_u_reduce(const double *__restrict__ part, int nsplit, int nf, int rs, double *__restrict__ U) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t m = (size_t)nf * rs;
    if (e >= m) return;
    const int i = (int)(e % (size_t)nf);
    double a = 0.0;
    for (int sp = (i / 64 * 64) / SCHUR_SK; sp < nsplit; sp++) a += part[(size_t)sp * m + e];
    U[e] = a;
}

// S = U(R, :) (rs x rs) with the identity below it: G (2 rs x rs, column-major,
// leading dimension ldg >= 2 rs), as the identity-row factor takes it.
__global__ void schur_s_aug(const double *__restrict__ U, int nf, const int32_t *__restrict__ R, int rs,
                            double *__restrict__ G, int ldg) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (a >= 2 * rs || b >= rs) return;
    G[(size_t)b * ldg + a] = (a < rs) ? U[(size_t)b * nf + R[a]] : (a - rs == b ? 1.0 : 0.0);
}

__global__ __launch_bounds__(256) void gather_from_factor(int n, int nf, const int32_t *__restrict__ map,
                                                          const double *__restrict__ zf, const double *__restrict__ r,
                                                          const double *__restrict__ A, int lda,
                                                          double *__restrict__ z, double *__restrict__ part) {
    __shared__ double wsum[2][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double p0 = 0.0, p1 = 0.0;
    if (i < n) {
        const int f = map[i];
        double z0, z1;
        if (f >= 0) {
            z0 = zf[f];
            z1 = zf[nf + f];
        } else {
            const double dii = A[(size_t)i * lda + i];
            z0 = r[i] / dii;
            z1 = r[n + i] / dii;
        }
        z[i] = z0;
        z[n + i] = z1;
        p0 = r[i] * z0;
        p1 = r[n + i] * z1;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        p0 += __shfl_xor(p0, o);
        p1 += __shfl_xor(p1, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wsum[0][wave] = p0;
        wsum[1][wave] = p1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int c = threadIdx.x;
        part[2 * blockIdx.x + c] = (wsum[c][0] + wsum[c][1]) + (wsum[c][2] + wsum[c][3]);
    }
}

// The failed-last-pivot result of the KKT system from its two SPD solves
// u = A^-1 r1, z = A^-1 y (X = [u; z]): s = -y^T z, b = (last - y^T u) / s^2,
// x = u - z b, x[n] = b.  One work-group; the two dot products are summed
// in a fixed order (strided per thread, then a tree).
__global__ __launch_bounds__(1024) void kkt_closed_form(int n, const double *__restrict__ X,
                                                        const double *__restrict__ ys, double last,
                                                        double *__restrict__ x) {
    __shared__ double sz[1024], su[1024];
    __shared__ double sb;
    const int t = threadIdx.x;
    double a = 0.0, c = 0.0;
    for (int i = t; i < n; i += 1024) {
        a = fma(ys[i], X[n + i], a);
        c = fma(ys[i], X[i], c);
    }
    sz[t] = a;
    su[t] = c;
    __syncthreads();
    for (int h = 512; h >= 1; h >>= 1) {
        if (t < h) {
            sz[t] += sz[t + h];
            su[t] += su[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double sv = 0.0 - sz[0];   // Schur value of the last pivot (H_nn = 0)
        sb = (last - su[0]) / (sv * sv);
        x[n] = sb;
    }
    __syncthreads();
    const double b = sb;
    for (int i = t; i < n; i += 1024) x[i] = fma(-X[n + i], b, X[i]);
}

struct PcgBufs {
    int cap = 0;
    double *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *w = nullptr;
    double *rf = nullptr, *zf = nullptr;
    double *part_a = nullptr, *part_b = nullptr, *part_c = nullptr, *rz = nullptr;
    double *y = nullptr;                       // [2][SCHUR_MAXR]: S^-1 zf(R) of the current application
    double *spart = nullptr, *sg = nullptr;    // split partials of U; S (tile, or with the identity below)
    size_t spart_n = 0;
};
thread_local PcgBufs g_pcg;

void pcg_reserve(int n) {
    if (g_pcg.cap >= n) return;
    double **vs[] = {&g_pcg.r, &g_pcg.p, &g_pcg.q, &g_pcg.z, &g_pcg.w, &g_pcg.rf, &g_pcg.zf};
    for (double **v : vs) {
        if (*v) irwls_dev_free(*v);
        *v = (double *)irwls_dev_alloc((size_t)2 * n * sizeof(double));
    }
    double **ps[] = {&g_pcg.part_a, &g_pcg.part_b, &g_pcg.part_c};
    for (double **v : ps) {
        if (*v) irwls_dev_free(*v);
        *v = (double *)irwls_dev_alloc((size_t)2 * (n + 1) * sizeof(double));
    }
    if (!g_pcg.rz) g_pcg.rz = (double *)irwls_dev_alloc(4 * sizeof(double));
    if (!g_pcg.y) g_pcg.y = (double *)irwls_dev_alloc(2 * SCHUR_MAXR * sizeof(double));
    g_pcg.cap = n;
}

// z = M^-1 r, with partials of r . z in part (one pair per block of 4 rows
// without a map, per block of 256 rows with one; npart returns the count).
// Without a map the factor's index set is the system's (nf = n).
struct Precond {
    const float *F;
    int ldf;
    const float *ZT;
    int ldz;
    int nf;
    const int32_t *map, *inv;
    const int32_t *R;       // removed rows' factor positions (rs of them, rs <= SCHUR_MAXR)
    int rs;
    const double *U, *Sinv; // B(:, R) and S^-1 (irwls_dev_pcg_schur_setup)
};

int precond(const Precond &M, int n, const double *A, int lda, const double *r, double *w, double *z,
            double *part) {
    const int nf = M.map ? M.nf : n;
    const int g = nf;   // one work-group, and one partial pair, per entry
    if (!M.map) {
        hipLaunchKernelGGL((coldot2<float, 1>), dim3(g), dim3(256), 0, stream(), M.F, M.ldf, (size_t)nf, nf, r, w,
                           (const double *)nullptr, (double *)nullptr);
        IRWLS_LAUNCH_CHECK();
        hipLaunchKernelGGL((coldot2<float, 2>), dim3(g), dim3(256), 0, stream(), M.ZT, M.ldz, (size_t)0, nf, w, z, r,
                           part);
        IRWLS_LAUNCH_CHECK();
        return g;
    }
    hipLaunchKernelGGL(scatter_to_factor, dim3((nf + 255) / 256), dim3(256), 0, stream(), n, nf, M.inv, r, g_pcg.rf);
    IRWLS_LAUNCH_CHECK();
    hipLaunchKernelGGL((coldot2<float, 1>), dim3(g), dim3(256), 0, stream(), M.F, M.ldf, (size_t)nf, nf, g_pcg.rf,
                       w, (const double *)nullptr, (double *)nullptr);
    IRWLS_LAUNCH_CHECK();
    hipLaunchKernelGGL((coldot2<float, 2>), dim3(g), dim3(256), 0, stream(), M.ZT, M.ldz, (size_t)0, nf, w, g_pcg.zf,
                       (const double *)nullptr, (double *)nullptr);
    IRWLS_LAUNCH_CHECK();
    if (M.rs > 0) {
        hipLaunchKernelGGL(schur_apply, dim3((M.rs + 3) / 4), dim3(256), 0, stream(), g_pcg.zf, nf, M.R, M.rs, M.Sinv,
                           g_pcg.y);
        IRWLS_LAUNCH_CHECK();
        hipLaunchKernelGGL(schur_correct, dim3((nf + 63) / 64), dim3(256), 0, stream(), g_pcg.zf, nf, M.U, M.rs,
                           (const double *)g_pcg.y);
        IRWLS_LAUNCH_CHECK();
    }
    const int gv = (n + 255) / 256;
    hipLaunchKernelGGL(gather_from_factor, dim3(gv), dim3(256), 0, stream(), n, nf, M.map, g_pcg.zf, r, A, lda, z,
                       part);
    IRWLS_LAUNCH_CHECK();
    return gv;
}

}  // namespace

extern "C" {

int irwls_dev_pcg_schur_setup(const float *ZT, int ldz, int nf, const int32_t *R, int rs, double *U, double *Sinv,
                              int check) {
    if (rs <= 0 || rs > SCHUR_MAXR || nf <= 0) {
        fprintf(stderr, "libirwls: irwls_dev_pcg_schur_setup needs 1 <= rs <= %d (rs=%d, nf=%d)\n", SCHUR_MAXR, rs,
                nf);
        exit(2);
    }
    const int nsplit = (nf + SCHUR_SK - 1) / SCHUR_SK;
    const size_t need = (size_t)nsplit * rs * nf;
    if (g_pcg.spart_n < need) {
        if (g_pcg.spart) irwls_dev_free(g_pcg.spart);
        g_pcg.spart = (double *)irwls_dev_alloc(need * sizeof(double));
        g_pcg.spart_n = need;
    }
    if (!g_pcg.sg) g_pcg.sg = (double *)irwls_dev_alloc((size_t)2 * SCHUR_MAXR * SCHUR_MAXR * sizeof(double));
    hipLaunchKernelGGL(schur_u_part, dim3((nf + 63) / 64, nsplit, (rs + 63) / 64), dim3(256), 0, stream(), ZT, ldz,
                       nf, R, rs, g_pcg.spart);
    IRWLS_LAUNCH_CHECK();
    const size_t m = (size_t)nf * rs;
    hipLaunchKernelGGL(schur_u_reduce, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream(), g_pcg.spart, nsplit,
                       nf, rs, U);
    IRWLS_LAUNCH_CHECK();
    // S^-1 = L_S^-T L_S^-1: from the tile inverse (rs <= 64, X = L_S^-1), or
    // from the identity rows of [S; I] factored (Z_S = L_S^-T)
    const int ldg = 2 * rs;
    hipLaunchKernelGGL(schur_s_aug, dim3((2 * rs + 255) / 256, rs), dim3(256), 0, stream(), U, nf, R, rs, g_pcg.sg,
                       ldg);
    IRWLS_LAUNCH_CHECK();
    int info;
    if (rs <= 64) {
        double *X = g_pcg.sg + (size_t)ldg * rs;   // 2 x 64 x 64 after G
        info = irwls_dev_tile_potrf_inv(g_pcg.sg, ldg, rs, X, check);
        irwls_dev_gemm(1, 0, rs, rs, rs, 1.0, X, 64, X, 64, 0.0, Sinv, rs);
    } else {
        info = irwls_dev_potrf_ident(g_pcg.sg, rs, ldg, check);
        irwls_dev_gemm(0, 1, rs, rs, rs, 1.0, g_pcg.sg + rs, ldg, g_pcg.sg + rs, ldg, 0.0, Sinv, rs);
    }
    return info;
}

int irwls_dev_pcg_spd2_mapped(const double *A, int n, int lda, const double *B, const float *F, int ldf,
                              const float *ZT, int ldz, int nf, const int32_t *map, const int32_t *inv, double tol,
                              int maxit, double *X, double *relres) {
    return irwls_dev_pcg_spd2_schur(A, n, lda, B, F, ldf, ZT, ldz, nf, map, inv, nullptr, 0, nullptr, nullptr, tol,
                                    maxit, X, relres);
}

int irwls_dev_pcg_spd2_schur(const double *A, int n, int lda, const double *B, const float *F, int ldf,
                             const float *ZT, int ldz, int nf, const int32_t *map, const int32_t *inv,
                             const int32_t *R, int rs, const double *U, const double *Sinv, double tol, int maxit,
                             double *X, double *relres) {
    if (n <= 0) return 0;
    pcg_reserve(n > nf ? n : nf);
    const Precond M = {F, ldf, ZT, ldz, nf, map, inv, R, map ? rs : 0, U, Sinv};
    const size_t vb = (size_t)2 * n * sizeof(double);
    const int gcol = n;                     // work-groups (and partial pairs) of the column products
    const int gvec = (n + 255) / 256;       // work-groups of the elementwise kernels
    TimerScope ts(IRWLS_T_PCG, 0.0, 0.0, false, true);
    // partials of |B|^2 sit after the residual partials in part_b (2 (n + 1)
    // entries >= 4 gvec) and travel with every residual read
    hipLaunchKernelGGL(sumsq2_parts, dim3(gvec), dim3(256), 0, stream(), n, B, g_pcg.part_b + 2 * gvec);
    IRWLS_LAUNCH_CHECK();
    irwls_dev_memset0(X, vb);
    irwls_dev_d2d(g_pcg.r, B, vb);
    int np = precond(M, n, A, lda, g_pcg.r, g_pcg.w, g_pcg.z, g_pcg.part_c);
    hipLaunchKernelGGL(update_p, dim3(gvec), dim3(256), 0, stream(), n, g_pcg.part_c, np, g_pcg.rz, 1, g_pcg.rz,
                       g_pcg.z, g_pcg.p);
    IRWLS_LAUNCH_CHECK();
    const size_t rrb = (size_t)4 * gvec * sizeof(double);
    // relative residuals from the partials read back for one iteration;
    // returns 1 when converged, -1 when diverged (or NaN), else 0
    auto judge = [&](const double *rr_host) {
        double rr[2] = {0.0, 0.0}, bn[2] = {0.0, 0.0};
        for (int g = 0; g < gvec; g++) {
            rr[0] += rr_host[2 * g];
            rr[1] += rr_host[2 * g + 1];
            bn[0] += rr_host[2 * gvec + 2 * g];
            bn[1] += rr_host[2 * gvec + 2 * g + 1];
        }
        for (int c = 0; c < 2; c++) relres[c] = (bn[c] > 0.0) ? sqrt(rr[c] / bn[c]) : 0.0;
        if (relres[0] <= tol && relres[1] <= tol) return 1;
        if (!(relres[0] < 1e30 && relres[1] < 1e30)) return -1;
        return 0;
    };
    int it = 0, done = 0;
    bool ap_queued = false;   // the product A p of this iteration was queued by the one before
    relres[0] = relres[1] = 1.0;
    auto queue_ap = [&]() {
        hipLaunchKernelGGL((coldot2<double, 0>), dim3(gcol), dim3(256), 0, stream(), A, lda, (size_t)0, n, g_pcg.p,
                           g_pcg.q, g_pcg.p, g_pcg.part_a);
        IRWLS_LAUNCH_CHECK();
    };
    for (it = 1; it <= maxit; it++) {
        double *rz_cur = g_pcg.rz + 2 * ((it - 1) & 1);
        double *rz_next = g_pcg.rz + 2 * (it & 1);
        if (!ap_queued) queue_ap();
        ap_queued = false;
        hipLaunchKernelGGL(update_xr, dim3(gvec), dim3(256), 0, stream(), n, g_pcg.part_a, gcol, rz_cur, g_pcg.p,
                           g_pcg.q, X, g_pcg.r, g_pcg.part_b);
        IRWLS_LAUNCH_CHECK();
        if (it == maxit) {
            readback_start(g_pcg.part_b, rrb);
            done = judge((const double *)readback_wait()) > 0;
            break;
        }
        // The residual of every other iteration is read back without
        // draining the stream: the preconditioner, the direction update and
        // the next product A p (none of which touches X) are queued before
        // the host waits for it, so a converged solve stops at the same
        // iteration, with the same X, as one that waits at once.
        const bool check = (it & 1) == 0;
        if (check) readback_start(g_pcg.part_b, rrb);
        np = precond(M, n, A, lda, g_pcg.r, g_pcg.w, g_pcg.z, g_pcg.part_c);
        hipLaunchKernelGGL(update_p, dim3(gvec), dim3(256), 0, stream(), n, g_pcg.part_c, np, rz_cur, 0, rz_next,
                           g_pcg.z, g_pcg.p);
        IRWLS_LAUNCH_CHECK();
        if (check) {
            queue_ap();
            ap_queued = true;
            const int v = judge((const double *)readback_wait());
            if (v != 0) {
                done = v > 0;
                break;
            }
        }
    }
    return done ? it : -it;
}

void irwls_dev_kkt_closed_form(int n, const double *X, const double *ys, double last, double *x) {
    if (n < 0) return;
    hipLaunchKernelGGL(kkt_closed_form, dim3(1), dim3(1024), 0, stream(), n, X, ys, last, x);
    IRWLS_LAUNCH_CHECK();
}

int irwls_dev_pcg_spd2(const double *A, int n, int lda, const double *B, const float *F, int ldf, const float *ZT,
                       int ldz, double tol, int maxit, double *X, double *relres) {
    return irwls_dev_pcg_spd2_mapped(A, n, lda, B, F, ldf, ZT, ldz, n, nullptr, nullptr, tol, maxit, X, relres);
}

void irwls_dev_release_pcg_buffers(void) {
    double **vs[] = {&g_pcg.r, &g_pcg.p, &g_pcg.q, &g_pcg.z, &g_pcg.w, &g_pcg.rf, &g_pcg.zf,
                     &g_pcg.part_a, &g_pcg.part_b, &g_pcg.part_c, &g_pcg.rz, &g_pcg.y, &g_pcg.spart, &g_pcg.sg};
    for (double **v : vs) {
        if (*v) irwls_dev_free(*v);
        *v = nullptr;
    }
    g_pcg.cap = 0;
    g_pcg.spart_n = 0;
}

}  // extern "C"
