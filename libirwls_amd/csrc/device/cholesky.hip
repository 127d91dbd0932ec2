// Blocked right-looking Cholesky and triangular solves for gfx950, FP64.
//
// Column-major lower factor in place, block size 64.  One step per block
// column k is two launches:
//   trsm:  the panel below the diagonal tile is multiplied by the inverse of
//          the factored tile (FP64 MFMA, v_mfma_f64_16x16x4f64);
//   syrk:  the trailing lower tiles take the rank-64 update (FP64 MFMA); the
//          work-group that owns the next diagonal tile then factors it and
//          forms its inverse in LDS (look-ahead), so the serial part of step
//          k+1 overlaps the trailing update of step k.
// Right-hand sides can ride along as extra rows below the matrix
// (rows n..nrows-1): the same steps then perform the forward sweep L z = r,
// so a solve only adds the backward sweep.
//
// A non-positive pivot ends the factorization as LAPACK potrf does: the Schur
// value stays on the diagonal, the failing index is reported and later steps
// are skipped; the right-hand-side rows are still solved against the factor
// as it stands, as LAPACK potrs does with it.  The reference relies on
// exactly this for the indefinite KKT matrix of full training
// (src/ParallelAlgorithms.c:277-279, 383-385, 441-445: LAPACK info is
// ignored and dpotrs uses the factor).
//
// No kernel here waits on another work-group: every dependency is a launch
// boundary on the library stream.
#include "common.h"

using namespace irwls;

namespace {

constexpr int NB = 64;
constexpr int BK = 16;
constexpr int LDP = BK + 1;

int *g_info = nullptr;     // [0] = potrf info

int *info_buf() {
    if (!g_info) g_info = (int *)irwls_dev_alloc(4 * sizeof(int));
    return g_info;
}

__device__ __forceinline__ double rdl(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ------------------------------------------------- diagonal tile factor ---
// Factor the lower kb x kb tile held in S (lower part) and form its inverse
// in Inv, with all 256 threads.
//   factor:  right-looking; per column one barrier publishes the scaled
//            column, one publishes the rank-1 update (independent LDS
//            updates, four threads per trailing row);
//   inverse: the four 16x16 diagonal sub-blocks are inverted by the four
//            waves in parallel (registers, one lane per column), then the
//            off-diagonal blocks follow block row by block row:
//            Linv_ij = -Linv_ii * sum_{j<=k<i} L_ik Linv_kj.
// Returns the failing column or -1.  After a failure the Schur value is left
// at L(j,j) (right-looking, so it is the fully updated pivot) and the inverse
// is formed from the tile as it stands.
__device__ int diag_factor_inverse(double (*S)[NB + 1], double (*Inv)[NB + 1], double *col, int kb) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int fail = -1;
    for (int j = 0; j < kb; j++) {
        const double ajj = S[j][j];
        if (!(ajj > 0.0)) {   // uniform
            fail = j;
            break;
        }
        const double ljj = sqrt(ajj);
        if (t < NB) col[t] = (t > j && t < kb) ? S[t][j] / ljj : ((t == j) ? ljj : 0.0);
        __syncthreads();
        if (t < NB && t >= j && t < kb) S[t][j] = col[t];
        const int i = j + 1 + (t >> 2);
        if (i < kb) {
            // all loads of this thread's share first, then the updates
            const double li = col[i];
            const int c0 = j + 1 + (t & 3);
            double sv[16], cv[16];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = c0 + 4 * q;
                const bool ok = c <= i;
                sv[q] = ok ? S[i][c] : 0.0;
                cv[q] = ok ? col[c] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = c0 + 4 * q;
                if (c <= i) S[i][c] = sv[q] - li * cv[q];
            }
        }
        __syncthreads();
    }
    // zero the strict upper triangle of the tile, as the factor is read by
    // block products below
    for (int e = t; e < NB * NB; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (c > r || r >= kb || c >= kb) S[r][c] = 0.0;
    }
    __syncthreads();
    // diagonal 16x16 sub-block inverses: wave w inverts block w, lane = column
    {
        const int b0 = wave * 16;
        const int c = lane;
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            double lrow[16];
#pragma unroll
            for (int k = 0; k <= i; k++) lrow[k] = S[b0 + i][b0 + k];
            double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; k++) s -= lrow[k] * x[k];
            x[i] = (c < 16 && i >= c && b0 + i < kb) ? s / lrow[i] : 0.0;
        }
        if (c < 16) {
#pragma unroll
            for (int i = 0; i < 16; i++) Inv[b0 + i][b0 + c] = x[i];
        }
    }
    // zero the strict upper blocks of Inv
    for (int e = t; e < NB * NB; e += 256) {
        const int r = e >> 6, c = e & 63;
        if ((c >> 4) > (r >> 4)) Inv[r][c] = 0.0;
    }
    __syncthreads();
    // off-diagonal blocks, block row by block row; thread -> (r, c) of a 16x16 block
    const int r = t >> 4, cc = t & 15;
#pragma unroll
    for (int bi = 1; bi < 4; bi++) {
        // T_j = sum_{k=j}^{bi-1} L_{bi,k} Linv_{k,j} for j < bi, kept in registers
        double tj[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int bj = 0; bj < bi; bj++) {
            double s = 0.0;
#pragma unroll
            for (int bk = bj; bk < bi; bk++) {
                double a[16], b[16];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    a[q] = S[bi * 16 + r][bk * 16 + q];
                    b[q] = Inv[bk * 16 + q][bj * 16 + cc];
                }
#pragma unroll
                for (int q = 0; q < 16; q++) s += a[q] * b[q];
            }
            tj[bj] = s;
        }
        __syncthreads();
        // stage T_j in the (still zero) strict upper blocks of Inv: block (bj, bi)
#pragma unroll
        for (int bj = 0; bj < bi; bj++) Inv[bj * 16 + r][bi * 16 + cc] = tj[bj];
        __syncthreads();
#pragma unroll
        for (int bj = 0; bj < bi; bj++) {
            double a[16], b[16];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                a[q] = Inv[bi * 16 + r][bi * 16 + q];
                b[q] = Inv[bj * 16 + q][bi * 16 + cc];
            }
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 16; q++) s += a[q] * b[q];
            tj[bj] = -s;
        }
        __syncthreads();
#pragma unroll
        for (int bj = 0; bj < bi; bj++) {
            Inv[bi * 16 + r][bj * 16 + cc] = (bi * 16 + r < kb) ? tj[bj] : 0.0;
            Inv[bj * 16 + r][bi * 16 + cc] = 0.0;
        }
        __syncthreads();
    }
    return fail;
}

// Write the factored tile (lower part) back and the inverse to Dinv
// (column-major 64x64, zero outside the lower kb x kb triangle).
__device__ void store_factor(double (*S)[NB + 1], double (*Inv)[NB + 1], double *A, int lda, int k0, int kb,
                             double *Dinv) {
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int i = e % NB, c = e / NB;
        if (i < kb && c < kb && c <= i) A[(size_t)(k0 + c) * lda + k0 + i] = S[i][c];
        Dinv[(size_t)c * NB + i] = (i < kb && c < kb && c <= i) ? Inv[i][c] : 0.0;
    }
}

__global__ __launch_bounds__(256) void diag_first_kernel(double *__restrict__ A, int lda, int kb,
                                                         int *__restrict__ info, double *__restrict__ Dinv) {
    __shared__ double S[NB][NB + 1];
    __shared__ double Inv[NB][NB + 1];
    __shared__ double col[NB];
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int i = e % NB, c = e / NB;
        S[i][c] = (i < kb && c < kb && c <= i) ? A[(size_t)c * lda + i] : 0.0;
    }
    __syncthreads();
    const int fail = diag_factor_inverse(S, Inv, col, kb);
    store_factor(S, Inv, A, lda, 0, kb, Dinv);
    if (threadIdx.x == 0 && fail >= 0) info[0] = fail + 1;
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

// acc[mb][nb] += sum_k P(i, k) Q(j, k) over k in [kc0, kc0+K), P and Q
// column-major with row offsets pr0 / qr0.
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
// Panel rows [k0+kb, nrows): P = A(:, k0:k0+kb) * Linv^T, in place.  After
// the factorization has failed in this block only right-hand-side rows
// (rows >= n) are solved; after a failure in an earlier block nothing is.
__global__ __launch_bounds__(256) void trsm_kernel(double *__restrict__ A, int n, int nrows, int lda, int k0, int kb,
                                                   const int *__restrict__ info, const double *__restrict__ Dinv) {
    __shared__ double As[1][NB][LDP];
    __shared__ double Bs[1][NB][LDP];
    const int inf = info[0];
    if (inf != 0 && inf <= k0) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = k0 + kb + blockIdx.x * NB;
    const int rmin = (inf != 0) ? n : 0;
    if (r0 + NB <= rmin) return;
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    // Q(j, k) = Linv(j, k) = Dinv[k*64 + j]
    mfma_cm(acc, As, Bs, A + (size_t)k0 * lda, lda, r0, nrows, Dinv, NB, 0, kb, 0, kb);
    for (int nb = 0; nb < 2; nb++) {
        const int c = wc * 32 + nb * 16 + (lane & 15);
        if (c >= kb) continue;
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) {
                const int i = r0 + wr * 32 + mb * 16 + (lane >> 4) + 4 * r;
                if (i < nrows && i >= rmin) A[(size_t)(k0 + c) * lda + i] = acc[mb][nb][r];
            }
    }
}

// ----------------------------------------------------------- syrk + next ---
// Trailing tiles: A(I, J) -= P_I P_J^T, rows [t0, nrows) x columns [t0, n),
// P = A(:, k0:k0+kb) (solved), t0 = k0 + kb; grid (column tiles, row tiles),
// tiles above the diagonal return at once.  Work-group (0,0) owns the next
// diagonal tile: after its update it factors and inverts it (look-ahead), so
// the next step needs no separate launch for that.
__global__ __launch_bounds__(256) void syrk_next_kernel(double *__restrict__ A, int n, int nrows, int lda, int k0,
                                                        int kb, int *__restrict__ info, double *__restrict__ Dnext) {
    __shared__ double As[1][NB][LDP];
    __shared__ double Bs[1][NB][LDP];
    __shared__ double S[NB][NB + 1];
    __shared__ double Inv[NB][NB + 1];
    __shared__ double colbuf[NB];
    const int tj = blockIdx.x, ti = blockIdx.y;
    const int t0 = k0 + kb;
    // stop only for a failure in blocks <= k (pivot index < t0); work-group
    // (0,0) may record a failure of block k+1 during this launch
    const int inf = info[0];
    if (ti < tj || (inf != 0 && inf <= t0)) return;
    const int rI = t0 + ti * NB, rJ = t0 + tj * NB;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++) acc[a][c] = (v4d){0.0, 0.0, 0.0, 0.0};
    mfma_cm(acc, As, Bs, A, lda, rI, nrows, A, lda, rJ, n, k0, kb);
    const bool diag = (ti == 0 && tj == 0);
    const int kb2 = (n - t0 < NB) ? n - t0 : NB;
    for (int nb = 0; nb < 2; nb++) {
        const int jl = wc * 32 + nb * 16 + (lane & 15);
        const int j = rJ + jl;
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) {
                const int il = wr * 32 + mb * 16 + (lane >> 4) + 4 * r;
                const int i = rI + il;
                const bool in_tile = diag && i < t0 + kb2;   // matrix rows of the next diagonal tile
                if (j < n && i < nrows && i >= j) {
                    const double v = A[(size_t)j * lda + i] - acc[mb][nb][r];
                    if (in_tile) S[il][jl] = v;
                    else A[(size_t)j * lda + i] = v;   // other tiles and right-hand-side rows
                }
                if (diag && !(in_tile && j < n && i >= j)) S[il][jl] = 0.0;
            }
    }
    if (!diag) return;
    __syncthreads();
    const int fail = diag_factor_inverse(S, Inv, colbuf, kb2);
    store_factor(S, Inv, A, lda, t0, kb2, Dnext);
    if (threadIdx.x == 0 && fail >= 0) info[0] = t0 + fail + 1;
}

// -------------------------------------------------------- triangular sweeps ---
// One launch per 64-row block w.  Every work-group first solves the diagonal
// block for block w of the solution from the (final) accumulator with wave 0
// and work-group 0 stores it in `out`; then the work-groups update the
// accumulators of the blocks still to come.  Block w of `acc` is only read in
// launch w, so no work-group reads what another one writes.

// forward: z_w = L_ww^-1 acc_w;  acc_r -= L(r, block w) z_w for rows r >= (w+1)*64
__global__ __launch_bounds__(256) void trsv_fwd_step(const double *__restrict__ L, int n, int lda, int w,
                                                     double *__restrict__ acc, double *__restrict__ out) {
    __shared__ double zs[NB];
    const int t = threadIdx.x, lane = t & 63;
    const int k0 = w * NB;
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    if (t < 64) {
        // lane i holds row i of L_ww
        double lrow[NB];
#pragma unroll
        for (int c = 0; c < NB; c++)
            lrow[c] = (lane < kb && c < kb && c <= lane) ? L[(size_t)(k0 + c) * lda + k0 + lane] : 0.0;
        double s = (lane < kb) ? acc[k0 + lane] : 0.0;
        double z = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < kb) {
                const double zj = rdl(s, j) / rdl(lrow[j], j);
                if (lane == j) z = zj;
                s -= lrow[j] * zj;   // rows below j; row j itself is finished
            }
        }
        zs[lane] = z;
        if (blockIdx.x == 0 && lane < kb) out[k0 + lane] = z;
    }
    __syncthreads();
    const int r = k0 + kb + blockIdx.x * 256 + t;
    if (r >= n) return;
    double v = 0.0;
    for (int c = 0; c < kb; c++) v += L[(size_t)(k0 + c) * lda + r] * zs[c];
    acc[r] -= v;
}

// backward: x_w = L_ww^-T acc_w;  acc_c -= sum_i L(w*64+i, c) x_w[i] for c < w*64
__global__ __launch_bounds__(256) void trsv_bwd_step(const double *__restrict__ L, int n, int lda, int w,
                                                     double *__restrict__ acc, double *__restrict__ out,
                                                     int cols_per_wave) {
    __shared__ double xs[NB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int k0 = w * NB;
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    if (t < 64) {
        // lane i holds column i of L_ww (row i of L_ww^T)
        double lcol[NB];
#pragma unroll
        for (int k = 0; k < NB; k++)
            lcol[k] = (lane < kb && k < kb && k >= lane) ? L[(size_t)(k0 + lane) * lda + k0 + k] : 0.0;
        double s = (lane < kb) ? acc[k0 + lane] : 0.0;
        double x = 0.0;
#pragma unroll
        for (int jj = NB - 1; jj >= 0; jj--) {
            if (jj < kb) {
                const double xj = rdl(s, jj) / rdl(lcol[jj], jj);
                if (lane == jj) x = xj;
                s -= lcol[jj] * xj;   // lanes i < jj take L(jj, i) x_jj
            }
        }
        xs[lane] = x;
        if (blockIdx.x == 0 && lane < kb) out[k0 + lane] = x;
    }
    __syncthreads();
    const double xl = xs[lane];
    const int c_begin = (blockIdx.x * 4 + wave) * cols_per_wave;
    for (int q = 0; q < cols_per_wave; q++) {
        const int c = c_begin + q;
        if (c >= k0) break;
        double v = (lane < kb) ? L[(size_t)c * lda + k0 + lane] * xl : 0.0;
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) acc[c] -= v;
    }
}

// right-hand-side row of an augmented matrix <-> contiguous vector
__global__ void gather_row(const double *__restrict__ A, int lda, int row, int n, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = A[(size_t)i * lda + row];
}

__global__ void scatter_row(double *__restrict__ A, int lda, int row, int n, const double *__restrict__ in) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(size_t)i * lda + row] = in[i];
}

double *g_dinv = nullptr;
int g_dinv_blocks = 0;

double *dinv_buf(int nblk) {
    if (g_dinv_blocks < nblk) {
        if (g_dinv) irwls_dev_free(g_dinv);
        g_dinv_blocks = nblk < 16 ? 16 : nblk;
        g_dinv = (double *)irwls_dev_alloc((size_t)g_dinv_blocks * NB * NB * sizeof(double));
    }
    return g_dinv;
}

void factor(double *A, int n, int nrows, int lda) {
    int *info = info_buf();
    irwls_dev_memset0(info, 4 * sizeof(int));
    const int nblk = (n + NB - 1) / NB;
    double *Dinv = dinv_buf(nblk);
    {
        const int kb = (n < NB) ? n : NB;
        TimerScope ts(IRWLS_T_POTF2, kb * (double)kb * kb / 3.0, 8.0 * (double)kb * kb);
        hipLaunchKernelGGL(diag_first_kernel, dim3(1), dim3(256), 0, stream(), A, lda, kb, info, Dinv);
        IRWLS_LAUNCH_CHECK();
    }
    for (int kbk = 0; kbk < nblk; kbk++) {
        const int k0 = kbk * NB;
        const int kb = (n - k0 < NB) ? (n - k0) : NB;
        const int prow = nrows - k0 - kb;
        if (prow > 0) {
            TimerScope ts(IRWLS_T_TRSM, 2.0 * prow * (double)kb * kb, 16.0 * prow * kb);
            hipLaunchKernelGGL(trsm_kernel, dim3((prow + NB - 1) / NB), dim3(256), 0, stream(), A, n, nrows, lda, k0,
                               kb, info, Dinv + (size_t)kbk * NB * NB);
            IRWLS_LAUNCH_CHECK();
        }
        const int rest = n - k0 - kb;
        if (rest <= 0) break;   // kernels return at once after a failure
        const int tc = (rest + NB - 1) / NB;
        const int tr = (nrows - k0 - kb + NB - 1) / NB;
        TimerScope ts(IRWLS_T_SYRK, (double)rest * rest * kb + 2.0 * (nrows - n) * rest * kb,
                      8.0 * (double)rest * rest);
        hipLaunchKernelGGL(syrk_next_kernel, dim3(tc, tr), dim3(256), 0, stream(), A, n, nrows, lda, k0, kb, info,
                           Dinv + (size_t)(kbk + 1) * NB * NB);
        IRWLS_LAUNCH_CHECK();
    }
}

// acc is consumed (left as scratch); the solution lands in out
void sweep_forward(const double *L, int n, int lda, double *acc, double *out) {
    const int nblk = (n + NB - 1) / NB;
    for (int w = 0; w < nblk; w++) {
        const int rows = n - (w + 1) * NB;
        const int g = rows > 0 ? (rows + 255) / 256 : 1;
        hipLaunchKernelGGL(trsv_fwd_step, dim3(g), dim3(256), 0, stream(), L, n, lda, w, acc, out);
        IRWLS_LAUNCH_CHECK();
    }
}

void sweep_backward(const double *L, int n, int lda, double *acc, double *out) {
    const int nblk = (n + NB - 1) / NB;
    const int cpw = 8;
    for (int w = nblk - 1; w >= 0; w--) {
        const int cols = w * NB;
        const int g = cols > 0 ? (cols + 4 * cpw - 1) / (4 * cpw) : 1;
        hipLaunchKernelGGL(trsv_bwd_step, dim3(g), dim3(256), 0, stream(), L, n, lda, w, acc, out, cpw);
        IRWLS_LAUNCH_CHECK();
    }
}

}  // namespace

extern "C" {

int irwls_dev_potrf(double *A, int n, int lda) {
    if (n <= 0) return 0;
    factor(A, n, n, lda);
    int info = 0;
    irwls_dev_d2h(&info, info_buf(), sizeof(int));
    return info;
}

void irwls_dev_potrs(const double *L, int n, int lda, double *b) {
    if (n <= 0) return;
    double *tmp = (double *)scratch(9, (size_t)n * sizeof(double));
    TimerScope ts(IRWLS_T_POTRS, 2.0 * (double)n * n, 8.0 * (double)n * n);
    sweep_forward(L, n, lda, b, tmp);     // z in tmp, b is scratch
    sweep_backward(L, n, lda, tmp, b);    // x in b
}

int irwls_dev_posv_aug(double *A, int n, int lda, const double *rhs, double *x) {
    if (n <= 0) return 0;
    if (lda < n + 1) {
        fprintf(stderr, "libirwls: irwls_dev_posv_aug needs lda >= n+1 (lda=%d, n=%d)\n", lda, n);
        exit(2);
    }
    hipLaunchKernelGGL(scatter_row, dim3((n + 255) / 256), dim3(256), 0, stream(), A, lda, n, n, rhs);
    IRWLS_LAUNCH_CHECK();
    factor(A, n, n + 1, lda);
    double *z = (double *)scratch(9, (size_t)n * sizeof(double));
    hipLaunchKernelGGL(gather_row, dim3((n + 255) / 256), dim3(256), 0, stream(), A, lda, n, n, z);
    IRWLS_LAUNCH_CHECK();
    {
        TimerScope ts(IRWLS_T_POTRS, 1.0 * (double)n * n, 4.0 * (double)n * n);
        sweep_backward(A, n, lda, z, x);
    }
    int info = 0;
    irwls_dev_d2h(&info, info_buf(), sizeof(int));
    return info;
}

}  // extern "C"
