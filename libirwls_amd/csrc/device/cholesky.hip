// Blocked right-looking Cholesky and triangular solves for gfx950, FP64.
//
// Column-major lower factor in place, block size 64:
//   potf2: one work-group factors the 64x64 diagonal block in LDS (wavefront
//          level, one column step per barrier pair) and forms its inverse;
//   trsm:  the panel below is multiplied by that inverse (FP64 MFMA);
//   syrk:  the trailing lower tiles take the rank-64 update (FP64 MFMA).
// A non-positive pivot ends the factorization as LAPACK potrf does: the Schur
// value stays on the diagonal, the failing index is reported, and later steps
// are skipped.  This is what the reference's solver relies on for the
// indefinite KKT matrix of full training (src/ParallelAlgorithms.c:277-279,
// 383-385, 441-445: LAPACK info is ignored and dpotrs uses the factor).
//
// potrs: forward and backward sweeps as one persistent launch each, one
// work-group per 64-row block, chained through per-block ready flags in
// global memory (agent-scope release/acquire).  Every wait is bounded: a
// wait that runs out sets an error word the host checks, so the kernel
// always terminates.
#include "common.h"

using namespace irwls;

namespace {

constexpr int NB = 64;
constexpr int BK = 16;
constexpr int LDP = BK + 1;

struct CholWork {
    double *Dinv = nullptr;     // per block 64x64 column-major inverse of L_kk
    int *info = nullptr;        // [0] = info, [1] = potrs wait error
    int *flags = nullptr;       // 2 * max blocks
    int max_blocks = 0;
    int epoch = 0;
};
CholWork g_cw;

void ensure_work(int nblk) {
    if (g_cw.max_blocks >= nblk && g_cw.info) return;
    if (g_cw.Dinv) {
        irwls_dev_free(g_cw.Dinv);
        irwls_dev_free(g_cw.flags);
    }
    int cap = nblk < 16 ? 16 : nblk;
    g_cw.Dinv = (double *)irwls_dev_alloc((size_t)cap * NB * NB * sizeof(double));
    g_cw.flags = (int *)irwls_dev_alloc((size_t)2 * cap * sizeof(int));
    if (!g_cw.info) g_cw.info = (int *)irwls_dev_alloc(4 * sizeof(int));
    g_cw.max_blocks = cap;
    g_cw.epoch = 0;
}

// ---------------------------------------------------------------- potf2 ---
__global__ __launch_bounds__(256) void potf2_kernel(double *__restrict__ A, int lda, int k0, int kb,
                                                    int *__restrict__ info, double *__restrict__ Dinv) {
    __shared__ double L[NB][NB];
    __shared__ double Inv[NB][NB];
    // After an earlier failure the factorization has stopped; the block is
    // only inverted as it stands, so the triangular solves use whatever the
    // lower triangle holds (as LAPACK potrs does).
    const bool stopped = info[0] != 0;
    const int t = threadIdx.x;
    for (int e = t; e < NB * NB; e += 256) {
        const int i = e % NB, c = e / NB;   // column-major walk: coalesced
        double v = 0.0;
        if (i < kb && c < kb && i >= c) v = A[(size_t)(k0 + c) * lda + k0 + i];
        L[i][c] = v;
    }
    __syncthreads();
    int fail = -1;
    if (!stopped) {
        for (int j = 0; j < kb; j++) {
            const double ajj = L[j][j];
            if (!(ajj > 0.0)) {   // uniform: every thread read the same value
                fail = j;
                break;
            }
            const double ljj = sqrt(ajj);
            __syncthreads();
            if (t == 0) L[j][j] = ljj;
            for (int i = j + 1 + t; i < kb; i += 256) L[i][j] = L[i][j] / ljj;
            __syncthreads();
            const int m = kb - j - 1;
            for (int e = t; e < m * m; e += 256) {
                const int i = j + 1 + e / m, c = j + 1 + e % m;
                if (c <= i) L[i][c] -= L[i][j] * L[c][j];
            }
            __syncthreads();
        }
        for (int e = t; e < NB * NB; e += 256) {
            const int i = e % NB, c = e / NB;
            if (i < kb && c < kb && i >= c) A[(size_t)(k0 + c) * lda + k0 + i] = L[i][c];
        }
        if (fail >= 0 && t == 0) info[0] = k0 + fail + 1;
    }
    // inverse of the lower triangular block, one column per thread
    if (t < NB) {
        for (int i = 0; i < NB; i++) {
            double x = 0.0;
            if (t < kb && i >= t && i < kb) {
                double s = (i == t) ? 1.0 : 0.0;
                for (int k = t; k < i; k++) s -= L[i][k] * Inv[k][t];
                x = s / L[i][i];
            }
            Inv[i][t] = x;
        }
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += 256) {
        const int i = e % NB, c = e / NB;
        Dinv[(size_t)c * NB + i] = Inv[i][c];
    }
}

// Loads a 64 (rows) x 16 (k) chunk of a column-major matrix:
// element (row, k) = M[(kc0 + k) * ld + r0 + row], zero outside [rows, kmax).
__device__ __forceinline__ void load_cm(double (*S)[LDP], const double *M, int ld, int r0, int rmax, int kc0,
                                        int kmax) {
    const int t = threadIdx.x;
    const int row = t & 63;
    const int kq = (t >> 6) * 4;
    const bool rok = (r0 + row) < rmax;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int k = kc0 + kq + q;
        double v = 0.0;
        if (rok && k < kmax) v = M[(size_t)k * ld + r0 + row];
        S[row][kq + q] = v;
    }
}

// acc[mb][nb] += sum_k P(i, k) Q(j, k) over k < K, P/Q as column-major
// matrices with row offsets pr0/qr0 and column offset kc0.
__device__ __forceinline__ void mfma_cm(v4d acc[2][2], double (*As)[NB][LDP], double (*Bs)[NB][LDP],
                                        const double *P, int ldp, int pr0, int prmax, const double *Q, int ldq,
                                        int qr0, int qrmax, int kc0, int K) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int nchunk = (K + BK - 1) / BK;
    for (int ch = 0; ch < nchunk; ch++) {
        load_cm(As[0], P, ldp, pr0, prmax, kc0 + ch * BK, kc0 + K);
        load_cm(Bs[0], Q, ldq, qr0, qrmax, kc0 + ch * BK, kc0 + K);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            const int kk = ks * 4 + (lane >> 4);
            double a0 = As[0][wr * 32 + (lane & 15)][kk];
            double a1 = As[0][wr * 32 + 16 + (lane & 15)][kk];
            double b0 = Bs[0][wc * 32 + (lane & 15)][kk];
            double b1 = Bs[0][wc * 32 + 16 + (lane & 15)][kk];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------- trsm ---
// Panel rows [k0+kb, n): P = A(:, k0:k0+kb) * inv(L_kk)^T, in place.
__global__ __launch_bounds__(256) void trsm_kernel(double *__restrict__ A, int n, int lda, int k0, int kb,
                                                   const int *__restrict__ info, const double *__restrict__ Dinv) {
    __shared__ double As[1][NB][LDP];
    __shared__ double Bs[1][NB][LDP];
    if (info[0] != 0) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = k0 + kb + blockIdx.x * NB;
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    // Q(j, k) = Linv(j, k) = Dinv[k*64 + j]  (column-major, ld 64)
    mfma_cm(acc, As, Bs, A + (size_t)k0 * lda, lda, r0, n, Dinv, NB, 0, kb, 0, kb);
    // all reads of the panel are done (barrier at the end of mfma_cm)
    for (int nb = 0; nb < 2; nb++) {
        const int c = wc * 32 + nb * 16 + (lane & 15);
        if (c >= kb) continue;
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) {
                const int i = r0 + wr * 32 + mb * 16 + (lane >> 4) + 4 * r;
                if (i < n) A[(size_t)(k0 + c) * lda + i] = acc[mb][nb][r];
            }
    }
}

// ----------------------------------------------------------------- syrk ---
// Trailing lower tiles: A(I, J) -= P_I P_J^T, P = A(:, k0:k0+kb).
__global__ __launch_bounds__(256) void syrk_kernel(double *__restrict__ A, int n, int lda, int k0, int kb,
                                                   const int *__restrict__ info) {
    __shared__ double As[1][NB][LDP];
    __shared__ double Bs[1][NB][LDP];
    if (info[0] != 0) return;
    // decode lower-triangular tile index: ti >= tj
    const int b = blockIdx.x;
    int ti = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ti++;
    while (ti * (ti + 1) / 2 > b) ti--;
    const int tj = b - ti * (ti + 1) / 2;
    const int t0 = k0 + kb;
    const int rI = t0 + ti * NB, rJ = t0 + tj * NB;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++) acc[a][c] = (v4d){0.0, 0.0, 0.0, 0.0};
    mfma_cm(acc, As, Bs, A, lda, rI, n, A, lda, rJ, n, k0, kb);
    for (int nb = 0; nb < 2; nb++) {
        const int j = rJ + wc * 32 + nb * 16 + (lane & 15);
        if (j >= n) continue;
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) {
                const int i = rI + wr * 32 + mb * 16 + (lane >> 4) + 4 * r;
                if (i < n && i >= j) A[(size_t)j * lda + i] -= acc[mb][nb][r];
            }
    }
}

// ---------------------------------------------------------------- potrs ---
constexpr long kSpinLimit = 1L << 22;   // ~0.1 s of s_sleep(1) polls

__device__ __forceinline__ void wait_flag(const int *flag, int epoch, int *err) {
    if (threadIdx.x == 0) {
        long spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            if (++spins > kSpinLimit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ __forceinline__ void post_flag(int *flag, int epoch) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Forward sweep L z = b, block w = blockIdx.x, in place on b.
__global__ __launch_bounds__(256) void trsv_fwd_kernel(const double *__restrict__ Lm, int n, int lda,
                                                       const double *__restrict__ Dinv, double *b, int *flags,
                                                       int epoch, int *err) {
    __shared__ double part[4][NB];
    __shared__ double accs[NB];
    const int w = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = w * NB + lane;
    double acc = 0.0;
    for (int j = 0; j < w; j++) {
        wait_flag(&flags[j], epoch, err);
        // this wave: columns j*64 + wave*16 .. +16
        for (int k = 0; k < 16; k++) {
            const int c = j * NB + wave * 16 + k;
            const double zc = __hip_atomic_load(&b[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (r < n) acc += Lm[(size_t)c * lda + r] * zc;
        }
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (threadIdx.x < NB) {
        const int i = threadIdx.x;
        const int gi = w * NB + i;
        const double rhs = (gi < n) ? b[gi] : 0.0;
        accs[i] = rhs - (((part[0][i] + part[1][i]) + part[2][i]) + part[3][i]);
    }
    __syncthreads();
    if (threadIdx.x < NB) {
        const int i = threadIdx.x;
        const int gi = w * NB + i;
        double z = 0.0;
        for (int k = 0; k <= i; k++) z += Dinv[(size_t)w * NB * NB + (size_t)k * NB + i] * accs[k];
        if (gi < n) b[gi] = z;
    }
    post_flag(&flags[w], epoch);
}

// Backward sweep L^T x = z, block w = nblk-1-blockIdx.x, in place on b.
__global__ __launch_bounds__(256) void trsv_bwd_kernel(const double *__restrict__ Lm, int n, int lda,
                                                       const double *__restrict__ Dinv, double *b, int *flags,
                                                       int epoch, int nblk, int *err) {
    __shared__ double part[4][16];
    __shared__ double accs[NB];
    const int w = nblk - 1 - blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // wave owns output rows w*64 + wave*16 + q, q < 16; lanes run over k
    double acc[16];
    for (int q = 0; q < 16; q++) acc[q] = 0.0;
    for (int j = w + 1; j < nblk; j++) {
        wait_flag(&flags[j], epoch, err);
        const int kr = j * NB + lane;
        const double xk = (kr < n) ? __hip_atomic_load(&b[kr], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        if (kr < n) {
            for (int q = 0; q < 16; q++) {
                const int c = w * NB + wave * 16 + q;
                if (c < n) acc[q] += Lm[(size_t)c * lda + kr] * xk;
            }
        }
    }
    for (int q = 0; q < 16; q++) {
        double v = acc[q];
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) part[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < NB) {
        const int i = threadIdx.x;
        const int gi = w * NB + i;
        const double rhs = (gi < n) ? b[gi] : 0.0;
        accs[i] = rhs - part[i >> 4][i & 15];
    }
    __syncthreads();
    if (threadIdx.x < NB) {
        const int i = threadIdx.x;
        const int gi = w * NB + i;
        // x_i = sum_k Linv(k, i) accs[k], Linv(k, i) = Dinv[i*64 + k], k >= i
        double x = 0.0;
        for (int k = i; k < NB; k++) x += Dinv[(size_t)w * NB * NB + (size_t)i * NB + k] * accs[k];
        if (gi < n) b[gi] = x;
    }
    post_flag(&flags[w], epoch);
}

}  // namespace

extern "C" {

int irwls_dev_potrf(double *A, int n, int lda) {
    if (n <= 0) return 0;
    const int nblk = (n + NB - 1) / NB;
    ensure_work(nblk);
    irwls_dev_memset0(g_cw.info, 4 * sizeof(int));
    for (int kbk = 0; kbk < nblk; kbk++) {
        const int k0 = kbk * NB;
        const int kb = (n - k0 < NB) ? (n - k0) : NB;
        {
            TimerScope ts(IRWLS_T_POTF2, kb * (double)kb * kb / 3.0, 16.0 * kb * kb);
            hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(256), 0, stream(), A, lda, k0, kb, g_cw.info,
                               g_cw.Dinv + (size_t)kbk * NB * NB);
            IRWLS_LAUNCH_CHECK();
        }
        const int rest = n - k0 - kb;
        if (rest <= 0) break;   // trsm/syrk return at once after a failure
        const int pt = (rest + NB - 1) / NB;
        {
            TimerScope ts(IRWLS_T_TRSM, 2.0 * rest * (double)kb * kb, 16.0 * rest * kb);
            hipLaunchKernelGGL(trsm_kernel, dim3(pt), dim3(256), 0, stream(), A, n, lda, k0, kb, g_cw.info,
                               g_cw.Dinv + (size_t)kbk * NB * NB);
            IRWLS_LAUNCH_CHECK();
        }
        {
            TimerScope ts(IRWLS_T_SYRK, (double)rest * rest * kb, 8.0 * (double)rest * rest);
            hipLaunchKernelGGL(syrk_kernel, dim3(pt * (pt + 1) / 2), dim3(256), 0, stream(), A, n, lda, k0, kb,
                               g_cw.info);
            IRWLS_LAUNCH_CHECK();
        }
    }
    int info = 0;
    irwls_dev_d2h(&info, g_cw.info, sizeof(int));
    return info;
}

void irwls_dev_potrs(const double *L, int n, int lda, double *b) {
    if (n <= 0) return;
    const int nblk = (n + NB - 1) / NB;
    ensure_work(nblk);
    // potrs uses the diagonal-block inverses formed by the last potrf of the
    // same matrix (also for blocks at or after a failed pivot).
    g_cw.epoch += 1;
    const int ep = g_cw.epoch;
    {
        TimerScope ts(IRWLS_T_POTRS, 2.0 * (double)n * n, 8.0 * (double)n * n);
        hipLaunchKernelGGL(trsv_fwd_kernel, dim3(nblk), dim3(256), 0, stream(), L, n, lda, g_cw.Dinv, b,
                           g_cw.flags, ep, g_cw.info + 1);
        IRWLS_LAUNCH_CHECK();
        hipLaunchKernelGGL(trsv_bwd_kernel, dim3(nblk), dim3(256), 0, stream(), L, n, lda, g_cw.Dinv, b,
                           g_cw.flags + g_cw.max_blocks, ep, nblk, g_cw.info + 1);
        IRWLS_LAUNCH_CHECK();
    }
    int err = 0;
    irwls_dev_d2h(&err, g_cw.info + 1, sizeof(int));
    if (err) {
        fprintf(stderr, "libirwls: triangular solve wait timed out\n");
        exit(2);
    }
}

}  // extern "C"
