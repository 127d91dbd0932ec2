// Budgeted-training kernels for gfx950 (SGMA basis selection and the
// parametric IRWLS), FP64 on the matrix cores where the work is GEMM-shaped.
//
//   gemm_f64       general C = alpha op(A) op(B) + beta C, column-major,
//                  64x64 tiles staged through LDS, v_mfma_f64_16x16x4f64;
//   sgma_gain      R = KSM - KSC Z fused with the per-candidate sum of R^2
//                  (src/budgeted-train.c:195-198): KSC streams once per step;
//   wgram          K1 += sum_i w_i phi_i phi_i^T over the active rows, split
//                  over the row range with a deterministic reduction
//                  (src/budgeted-train.c:401-416);
//   resid_rows     e = y - Phi beta, one wave per row (src/budgeted-train.c:433-443);
//   irls_weights   D_i = C/(y_i e_i) clipped to M, 0 on the wrong side, and the
//                  active-row compaction in row order (src/budgeted-train.c:448-465).
#include "common.h"

using namespace irwls;

namespace {

constexpr int T = 64;
constexpr int BK = 16;
constexpr int LDP = BK + 1;

// A 64 x BK chunk of an operand, element (row, k) = M[(kc0 + k) * ld + r0 +
// row] (k-outer layout) or M[(r0 + row) * ld + kc0 + k] (ROUTER, row-outer),
// zero past rmax rows or kmax: fetched into registers, then stored to LDS,
// so that the next chunk's loads fly while the current one multiplies.
template <bool ROUTER>
__device__ __forceinline__ void fetch_chunk(double v[4], const double *M, int ld, int r0, int rmax, int kc0, int kmax) {
    const int t = threadIdx.x;
    const int row = ROUTER ? t >> 2 : t & 63;
    const int kq = ROUTER ? (t & 3) * 4 : (t >> 6) * 4;
    const bool rok = (r0 + row) < rmax;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int k = kc0 + kq + q;
        v[q] = (rok && k < kmax) ? (ROUTER ? M[(size_t)(r0 + row) * ld + k] : M[(size_t)k * ld + r0 + row]) : 0.0;
    }
}
template <bool ROUTER>
__device__ __forceinline__ void put_chunk(double (*S)[LDP], const double v[4]) {
    const int t = threadIdx.x;
    const int row = ROUTER ? t >> 2 : t & 63;
    const int kq = ROUTER ? (t & 3) * 4 : (t >> 6) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) S[row][kq + q] = v[q];
}

__device__ __forceinline__ void mfma_chunk(v4d acc[2][2], double (*As)[LDP], double (*Bs)[LDP]) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
    for (int ks = 0; ks < BK / 4; ks++) {
        const int kk = ks * 4 + (lane >> 4);
        const double a0 = As[wr * 32 + (lane & 15)][kk];
        const double a1 = As[wr * 32 + 16 + (lane & 15)][kk];
        const double b0 = Bs[wc * 32 + (lane & 15)][kk];
        const double b1 = Bs[wc * 32 + 16 + (lane & 15)][kk];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// tile element owned by (mb, nb, r) of this lane
__device__ __forceinline__ int tile_row(int mb, int r) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave >> 1) * 32 + mb * 16 + (lane >> 4) + 4 * r;
}
__device__ __forceinline__ int tile_col(int nb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave & 1) * 32 + nb * 16 + (lane & 15);
}

// ------------------------------------------------------------------- gemm ---
// Split over k along z when the tile grid is too small to fill the device:
// work-group z sums k in [z * kper, (z + 1) * kper) and, with more than one
// slice, stores its raw tile to part[z][j][i] (m x n, column-major) for
// gemm_splitk_finish, which adds the slices in slice order.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(int m, int n, int k, double alpha, const double *__restrict__ A,
                                                       int lda, const double *__restrict__ B, int ldb, double beta,
                                                       double *__restrict__ C, int ldc, int kper,
                                                       double *__restrict__ part) {
    __shared__ double As[2][T][LDP];
    __shared__ double Bs[2][T][LDP];
    const int r0 = blockIdx.x * T, c0 = blockIdx.y * T;
    const int kb = blockIdx.z * kper, ke = min(k, kb + kper);
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    // op(A)(i, p): !TA -> A[i + p*lda] (k-outer); TA -> A[p + i*lda] (row-outer)
    // op(B)(p, j) as (row j, k p): !TB -> B[p + j*ldb] (row-outer); TB -> B[j + p*ldb] (k-outer)
    // The chunks are double-buffered in LDS: the next chunk is fetched into
    // registers before the current one is multiplied, one barrier per chunk.
    double va[4], vb[4];
    if (kb < ke) {
        fetch_chunk<TA>(va, A, lda, r0, m, kb, ke);
        fetch_chunk<!TB>(vb, B, ldb, c0, n, kb, ke);
        put_chunk<TA>(As[0], va);
        put_chunk<!TB>(Bs[0], vb);
    }
    __syncthreads();
    int buf = 0;
    for (int kc = kb; kc < ke; kc += BK) {
        const bool more = kc + BK < ke;
        if (more) {
            fetch_chunk<TA>(va, A, lda, r0, m, kc + BK, ke);
            fetch_chunk<!TB>(vb, B, ldb, c0, n, kc + BK, ke);
        }
        mfma_chunk(acc, As[buf], Bs[buf]);
        if (more) {
            put_chunk<TA>(As[buf ^ 1], va);
            put_chunk<!TB>(Bs[buf ^ 1], vb);
        }
        __syncthreads();
        buf ^= 1;
    }
    for (int nb = 0; nb < 2; nb++) {
        const int j = c0 + tile_col(nb);
        if (j >= n) continue;
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) {
                const int i = r0 + tile_row(mb, r);
                if (i >= m) continue;
                if (gridDim.z > 1) {
                    part[((size_t)blockIdx.z * n + j) * m + i] = acc[mb][nb][r];
                } else {
                    double *p = C + (size_t)j * ldc + i;
                    *p = (beta == 0.0) ? alpha * acc[mb][nb][r] : alpha * acc[mb][nb][r] + beta * (*p);
                }
            }
    }
}

__global__ void gemm_splitk_finish(int m, int n, int nz, const double *__restrict__ part, double alpha, double beta,
                                   double *__restrict__ C, int ldc) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)m * n) return;
    const int i = (int)(e % m), j = (int)(e / m);
    double s = 0.0;
    for (int z = 0; z < nz; z++) s += part[(size_t)z * m * n + e];
    double *p = C + (size_t)j * ldc + i;
    *p = (beta == 0.0) ? alpha * s : alpha * s + beta * (*p);
}

// -------------------------------------------------------------- sgma_gain ---
// R(i, c) = KSM[i*ldm + c] - sum_p KSC[i + p*n] Z[p + c*ldz] over a tile of
// GR = 128 rows x 64 candidates; part[rowtile * ncand + c] = sum over the
// tile's rows of R(i, c)^2.  Each wave owns 32 rows and every candidate of
// the tile (2 x 4 blocks of v_mfma_f64_16x16x4f64).  Its KSC operands come
// from HBM straight into registers one 16-k chunk ahead of the products
// (KSC is k-outer, so one block's 16 rows at one k are 128 contiguous
// bytes); the Z chunk, shared by the four waves, is double-buffered in LDS,
// one barrier per chunk.  KSC streams once per column tile: 16 flops per
// byte, at the FP64 matrix-core ridge of the part.
constexpr int GR = 128, GK = 16;

__device__ __forceinline__ void gain_load_a(double a[4][2], const double *__restrict__ KSC, int n, int s, int kc,
                                            int ra0, bool ok0, bool ok1) {
    const int kq = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int k = kc + 4 * u + kq;
        const bool kok = k < s;
        a[u][0] = (ok0 && kok) ? KSC[(size_t)k * n + ra0] : 0.0;
        a[u][1] = (ok1 && kok) ? KSC[(size_t)k * n + ra0 + 16] : 0.0;
    }
}

template <int TC>
__global__ __launch_bounds__(256) void sgma_gain_kernel(int n, int s, int ncand, const double *__restrict__ KSC,
                                                        const double *__restrict__ Z, int ldz,
                                                        const double *__restrict__ KSM, int ldm,
                                                        double *__restrict__ part, int kchunk,
                                                        double *__restrict__ rpart) {
    __shared__ double Zs[2][GK][TC];
    __shared__ double red[4][TC];
    // k range of this work-group: all of [0, s) unless split (rpart != NULL)
    const int kb = blockIdx.z * kchunk, ke = min(s, kb + kchunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int r0 = blockIdx.x * GR + wave * 32;
    const int c0 = blockIdx.y * TC;
    const int ra0 = r0 + li;
    const bool ok0 = ra0 < n, ok1 = ra0 + 16 < n;
    // Z staging: thread t carries candidate c0 + t / (GK / ZK), k = ZK (t % (GK / ZK)) .. +ZK-1 of a chunk
    constexpr int ZK = GK * TC / 256, NB = TC / 16;
    const int zc = threadIdx.x / (GK / ZK), zk = (threadIdx.x % (GK / ZK)) * ZK;
    const bool zok = c0 + zc < ncand;
    const double *zp = Z + (size_t)(zok ? c0 + zc : 0) * ldz;
    double z[ZK];
#pragma unroll
    for (int q = 0; q < ZK; q++) z[q] = (zok && kb + zk + q < ke) ? zp[kb + zk + q] : 0.0;
#pragma unroll
    for (int q = 0; q < ZK; q++) Zs[0][zk + q][zc] = z[q];
    double a[4][2], an[4][2];
    gain_load_a(a, KSC, n, ke, kb, ra0, ok0, ok1);
    v4d acc[2][NB];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mb][nb] = (v4d){0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    int buf = 0;
    for (int kc = kb; kc < ke; kc += GK) {
        const bool more = kc + GK < ke;
        if (more) {
            gain_load_a(an, KSC, n, ke, kc + GK, ra0, ok0, ok1);
#pragma unroll
            for (int q = 0; q < ZK; q++) {
                const int k = kc + GK + zk + q;
                z[q] = (zok && k < ke) ? zp[k] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            double bv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) bv[nb] = Zs[buf][4 * u + kq][nb * 16 + li];
#pragma unroll
            for (int mb = 0; mb < 2; mb++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][mb], bv[nb], acc[mb][nb], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < ZK; q++) Zs[buf ^ 1][zk + q][zc] = z[q];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                a[u][0] = an[u][0];
                a[u][1] = an[u][1];
            }
        }
        __syncthreads();
        buf ^= 1;
    }
    // accumulator register r of lane l: (row l/16 + 4r, candidate l%16) of its block
    if (rpart) {   // split k: this range's products, summed by gain_split_finish
        double *rp = rpart + (size_t)blockIdx.z * n * ncand;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int c = c0 + nb * 16 + li;
#pragma unroll
            for (int mb = 0; mb < 2; mb++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = r0 + mb * 16 + kq + 4 * r;
                    if (i < n && c < ncand) rp[(size_t)i * ncand + c] = acc[mb][nb][r];
                }
        }
        return;
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const int c = c0 + nb * 16 + li;
        double q = 0.0;
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = r0 + mb * 16 + kq + 4 * r;
                if (i < n && c < ncand) {
                    const double v = KSM[(size_t)i * ldm + c] - acc[mb][nb][r];
                    q += v * v;
                }
            }
        // lanes l, l^16, l^32, l^48 hold the same candidate
        q += __shfl_xor(q, 16);
        q += __shfl_xor(q, 32);
        if (lane < 16) red[wave][nb * 16 + lane] = q;
    }
    __syncthreads();
    if (threadIdx.x < TC) {
        const int c = c0 + threadIdx.x;
        if (c < ncand)
            part[(size_t)blockIdx.x * ncand + c] =
                (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
    }
}

// Gain sums for at most 16 candidates: one 16 x 16 MFMA block per wave (16
// rows, every candidate); the work-group covers 64 rows.  KSC operands come
// straight from HBM into registers one 16-k chunk ahead of the products
// (k-outer, so 16 consecutive rows per k); the 16 x 16 chunk of Z, shared by
// the four waves, is double-buffered in LDS (one load per thread, one
// barrier per chunk), so the Z reads cost no per-candidate cache lines in
// the vector memory path.  Past k = s the operands are zero, which adds
// exact zeros.  Bound by the one pass over KSC.
__global__ __launch_bounds__(256) void sgma_gain_narrow_kernel(int n, int s, int ncand, const double *__restrict__ KSC,
                                                               const double *__restrict__ Z, int ldz,
                                                               const double *__restrict__ KSM, int ldm,
                                                               double *__restrict__ part, int kchunk,
                                                               double *__restrict__ rpart) {
    __shared__ double Zs[2][16][16];     // [k][candidate]
    __shared__ double red[4][16];
    // k range of this work-group: all of [0, s) unless split (rpart != NULL)
    const int kb = blockIdx.z * kchunk, ke = min(s, kb + kchunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.x * T + wave * 16;
    const int ra = row0 + (lane & 15);        // A operand row
    const int cb = lane & 15;                 // B operand column (candidate)
    const int kq = lane >> 4;
    const bool rok = ra < n, cok = cb < ncand;
    const double *ap = KSC + (rok ? ra : 0);
    // Z staging: thread t carries candidate t / 16, k = t % 16 of a chunk
    const int zc = threadIdx.x >> 4, zk = threadIdx.x & 15;
    const bool zok = zc < ncand;
    const double *zp = Z + (size_t)(zok ? zc : 0) * ldz;
    double a[4], an[4];
    double zv = (zok && kb + zk < ke) ? zp[kb + zk] : 0.0;
    Zs[0][zk][zc] = zv;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int k = kb + 4 * u + kq;
        a[u] = (rok && k < ke) ? ap[(size_t)k * n] : 0.0;
    }
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += 16) {
        const bool more = k0 + 16 < ke;
        if (more) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + 16 + 4 * u + kq;
                an[u] = (rok && k < ke) ? ap[(size_t)k * n] : 0.0;
            }
            const int k = k0 + 16 + zk;
            zv = (zok && k < ke) ? zp[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], Zs[buf][4 * u + kq][cb], acc, 0, 0, 0);
        if (more) {
            Zs[buf ^ 1][zk][zc] = zv;
#pragma unroll
            for (int u = 0; u < 4; u++) a[u] = an[u];
        }
        __syncthreads();
        buf ^= 1;
    }
    // accumulator register r of lane l: (row (l/16) + 4r, candidate l%16)
    if (rpart) {   // split k: this range's products, summed by gain_split_finish
        double *rp = rpart + (size_t)blockIdx.z * n * ncand;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = row0 + (lane >> 4) + 4 * r;
            if (i < n && cok) rp[(size_t)i * ncand + cb] = acc[r];
        }
        return;
    }
    double q = 0.0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = row0 + (lane >> 4) + 4 * r;
        if (i < n && cok) {
            const double v = KSM[(size_t)i * ldm + cb] - acc[r];
            q += v * v;
        }
    }
    q += __shfl_xor(q, 16);
    q += __shfl_xor(q, 32);
    if (lane < 16) red[wave][lane] = q;
    __syncthreads();
    if (threadIdx.x < 16 && threadIdx.x < ncand) {
        const int c = threadIdx.x;
        part[(size_t)blockIdx.x * ncand + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
    }
}

// Split-k gains, finish: part[b][c] = sum over the rb rows i of row block b
// of (KSM[i][c] - sum_z rpart[z][i][c])^2, the k ranges added in order
// z = 0, 1, ...  One wave per block: lane l takes candidate l % C of each
// group of C (the power of two >= ncand, at most 64) and rows l / C + j (64 /
// C) of the block; the row lanes of a candidate add by a fixed butterfly.
// rb = gain_split_rows(ncand): 4 rows from 17 candidates on, 16 below, so
// that a lane adds at most 4 rows.
__device__ __host__ inline int gain_split_c(int ncand) {
    int C = 1;
    while (C < ncand && C < 64) C <<= 1;
    return C;
}
__device__ __host__ inline int gain_split_rows(int ncand) { return gain_split_c(ncand) >= 32 ? 4 : 16; }

__global__ __launch_bounds__(256) void gain_split_finish(int n, int ncand, int nz, const double *__restrict__ rpart,
                                                         const double *__restrict__ KSM, int ldm,
                                                         double *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int rb = gain_split_rows(ncand);
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int r0 = b * rb;
    if (r0 >= n) return;   // whole waves; the kernel has no barrier
    const int C = gain_split_c(ncand);
    const int cl = lane & (C - 1), rl = lane / C, RL = 64 / C;
    const int r1 = min(n, r0 + rb);
    const size_t zs = (size_t)n * ncand;
    for (int c0 = 0; c0 < ncand; c0 += C) {
        const int c = c0 + cl;
        double q = 0.0;
        if (c < ncand) {
#pragma unroll 4
            for (int i = r0 + rl; i < r1; i += RL) {
                const double *rp = rpart + (size_t)i * ncand + c;
                double v = rp[0];
#pragma unroll
                for (int z = 1; z < 8; z++)
                    if (z < nz) v += rp[z * zs];
                const double d = KSM[(size_t)i * ldm + c] - v;
                q += d * d;
            }
        }
        for (int o = C; o < 64; o <<= 1) q += __shfl_xor(q, o);
        if (rl == 0 && c < ncand) part[(size_t)b * ncand + c] = q;
    }
}

// fixed-order column sums of the partials: one work-group per column, each
// thread a strided subset in order, then a tree in a fixed order (the same
// sum on every run)
__global__ __launch_bounds__(256) void col_sum_kernel(const double *__restrict__ part, int nparts, int ncol,
                                                      double *__restrict__ out) {
    __shared__ double sh[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += 256) s += part[(size_t)p * ncol + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// eta_c = 1 - sum_e KNC[c*ldk + e] Z[e + c*ldz], one wave per candidate
__global__ void sgma_eta_kernel(int s, int ncand, const double *__restrict__ KNC, int ldk,
                                const double *__restrict__ Z, int ldz, double *__restrict__ eta) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= ncand) return;
    double v = 0.0;
    for (int e = lane; e < s; e += 64) v += KNC[(size_t)c * ldk + e] * Z[e + (size_t)c * ldz];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) eta[c] = 1.0 - v;
}

// dst[i] = src[i*ld + col]
__global__ void gather_col_kernel(const double *__restrict__ src, int ld, int col, int n, double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(size_t)i * ld + col];
}

// M[c*ld + row] = v[c] for c < count
__global__ void scatter_row_kernel(double *__restrict__ M, int ld, int row, int count, const double *__restrict__ v) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < count) M[(size_t)c * ld + row] = v[c];
}

// Row s of L and L^-1 of K_C + jitter (src/budgeted-train.c:247-279) from
// L2 = L^-1 k and IL2 = L2^T L^-1: one thread forms
// L3 = sqrt(kcc + 1e-5 - sum_i L2_i^2) with the sum taken in index order and
// every product and difference rounded on its own (no fused multiply-add),
// and IL3 = 1 / L3; then L(s, 0..s) = (L2, L3), L^-1(s, 0..s) = (-IL3 IL2, IL3).
// The squares are formed by all threads into LDS (LROW_MAX at a time); the
// one thread's chain is then only the differences, read sixteen at a time.
constexpr int LROW_MAX = 4096;
__global__ __launch_bounds__(256) void sgma_lrow_kernel(int s, int ld, const double *__restrict__ kcc,
                                                        const double *__restrict__ L2,
                                                        const double *__restrict__ IL2, double *__restrict__ L,
                                                        double *__restrict__ Li) {
    __shared__ double sq[LROW_MAX];
    __shared__ double sh[2];
    double acc = kcc[0] + 0.00001;   // carried by thread 0
    for (int i0 = 0; i0 < s; i0 += LROW_MAX) {
        const int m = (s - i0 < LROW_MAX) ? s - i0 : LROW_MAX;
        __syncthreads();   // the previous segment has been read
        for (int i = threadIdx.x; i < m; i += blockDim.x) sq[i] = __dmul_rn(L2[i0 + i], L2[i0 + i]);
        __syncthreads();
        if (threadIdx.x == 0) {
            int i = 0;
            for (; i + 16 <= m; i += 16) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; u++) v[u] = sq[i + u];
#pragma unroll
                for (int u = 0; u < 16; u++) acc = __dsub_rn(acc, v[u]);
            }
            for (; i < m; i++) acc = __dsub_rn(acc, sq[i]);
        }
    }
    if (threadIdx.x == 0) {
        const double r = sqrt(acc);
        sh[0] = r;
        sh[1] = 1.0 / r;
    }
    __syncthreads();
    const double L3 = sh[0], IL3 = sh[1];
    for (int c = threadIdx.x; c <= s; c += blockDim.x) {
        L[(size_t)c * ld + s] = (c < s) ? L2[c] : L3;
        Li[(size_t)c * ld + s] = (c < s) ? __dmul_rn(-IL3, IL2[c]) : IL3;
    }
}

__global__ void sgma_pick_kernel(int nc, const double *__restrict__ q, const double *__restrict__ eta,
                                 const int32_t *__restrict__ cand, int s, int n, const double *__restrict__ y,
                                 int32_t *__restrict__ cent, int32_t *__restrict__ st, double *__restrict__ value) {
    if (threadIdx.x != 0) return;
    double bestv = 0.0;
    int best = 0;
    for (int i = 0; i < nc; i++) {
        const double g = (eta[i] > 0.0) ? __dmul_rn(1.0 / eta[i], q[i]) : 0.0;
        if (i == 0 || g > bestv) {
            bestv = g;
            best = i;
        }
    }
    st[0] = best;
    value[0] = bestv;
    const int c = cand[best];
    cent[s] = c;
    if (c == n || (c < n && y[c] > 0.0)) st[1] += 1;
}

__global__ void gather_col_at_kernel(const double *__restrict__ src, int ld, const int32_t *__restrict__ col, int n,
                                     double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(size_t)i * ld + col[0]];
}

__global__ void copy_row_at_kernel(const double *__restrict__ src, int ld, const int32_t *__restrict__ row, int count,
                                   double *__restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < count) dst[j] = src[(size_t)row[0] * ld + j];
}

// One wave: lane l holds entries l + 64 u (u < 16, P <= 1024); round r takes
// the first untaken entry if its score is NaN, else the largest score among
// the untaken non-NaN ones with the smallest index, which is what a serial
// scan `if (!taken[i] && (bi < 0 || score[i] > score[bi])) bi = i` picks.
__global__ __launch_bounds__(64) void sgma_topk_kernel(int P, const float *__restrict__ score, int k,
                                                       const int32_t *__restrict__ cidx, int32_t *__restrict__ out,
                                                       int32_t *__restrict__ pos) {
    const int lane = threadIdx.x;
    const int none = 1 << 30;
    float sc[16];
#pragma unroll
    for (int u = 0; u < 16; u++) sc[u] = (lane + 64 * u < P) ? score[lane + 64 * u] : 0.0f;
    unsigned taken = 0u;
    for (int r = 0; r < k; r++) {
        int first = none, bi = none;
        float bv = 0.0f;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int i = lane + 64 * u;
            if (i >= P || ((taken >> u) & 1u)) continue;
            if (first == none) first = i;
            const float v = sc[u];
            if (v == v && (bi == none || v > bv)) {
                bv = v;
                bi = i;
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            const int f2 = __shfl_xor(first, o);
            const float v2 = __shfl_xor(bv, o);
            const int i2 = __shfl_xor(bi, o);
            first = f2 < first ? f2 : first;
            if (i2 != none && (bi == none || v2 > bv || (v2 == bv && i2 < bi))) {
                bv = v2;
                bi = i2;
            }
        }
        int pick = bi;
        if (first != none) {
            const float f = __shfl(sc[0], first & 63);   // sc[first >> 6] of lane first & 63
            float fv = f;
#pragma unroll
            for (int u = 1; u < 16; u++) {
                const float g = __shfl(sc[u], first & 63);
                if ((first >> 6) == u) fv = g;
            }
            if (fv != fv || pick == none) pick = first;
        }
        if (pick == none) pick = 0;   // k > P: not reached (the caller keeps k <= P)
        if (lane == (pick & 63)) taken |= 1u << (pick >> 6);
        if (lane == 0) {
            out[r] = cidx[pick];
            if (pos) pos[r] = pick;
        }
    }
}

// Shortlist operands from the pool's: row r of KNC2 and column r of Z2 (s
// entries each, leading dimension ld) and eta2[r] are those of pool position
// pos[r].  One work-group per shortlisted candidate.
__global__ __launch_bounds__(256) void gather_shortlist_kernel(int s, int ld, const int32_t *__restrict__ pos,
                                                               const double *__restrict__ KNC,
                                                               const double *__restrict__ Z,
                                                               const double *__restrict__ eta,
                                                               double *__restrict__ KNC2, double *__restrict__ Z2,
                                                               double *__restrict__ eta2) {
    const int r = blockIdx.x;
    const int p = pos[r];
    for (int e = threadIdx.x; e < s; e += 256) {
        KNC2[(size_t)r * ld + e] = KNC[(size_t)p * ld + e];
        Z2[(size_t)r * ld + e] = Z[(size_t)p * ld + e];
    }
    if (threadIdx.x == 0) eta2[r] = eta[p];
}

// ----------------------------------------------------------------- wgram ---
// Partial Gram matrices over row chunks: part[split][c*B + r] (lower tiles,
// mirrored) = sum_{q in chunk} w_q Phi[a_q][r] Phi[a_q][c] with
// w_q = sqrt(D[a_q])^2 taken as the product of the two scaled rows, as the
// reference scales the rows by sqrt(D) before the product; a = active list
// (NULL = all rows, weights 1).  Phi is row-major n x B.
//
// A work-group takes a 128 x 128 lower tile (ti >= tj) and one row chunk;
// its waves own 64 x 64 quarters (4 x 4 blocks of v_mfma_f64_16x16x4f64).
// The next 16 rows are loaded into registers (each thread 8 consecutive
// centroids of one row, both tile sides) while the current 16 are
// multiplied out of LDS, double-buffered, one barrier per 16 rows.  With
// xcd_map, work-group b runs on XCD b % 8 and takes split
// (b % 8) + 8 * (b / (8 * ntiles)), so the tiles that read one row chunk
// share that XCD's L2 and the chunk streams from HBM about once.
constexpr int WT = 128, WK = 16, WLP = WK + 1;

__global__ __launch_bounds__(256) void wgram_kernel(int B, const double *__restrict__ Phi, const int32_t *__restrict__ act,
                                                    const double *__restrict__ sw, int nact, int chunk, int ntiles,
                                                    int xcd_map, double *__restrict__ part) {
    __shared__ double Sa[2][WT][WLP];
    __shared__ double Sb[2][WT][WLP];
    const int bid = blockIdx.x;
    int tile, split;
    if (xcd_map) {
        const int r = bid >> 3;
        tile = r % ntiles;
        split = (bid & 7) + 8 * (r / ntiles);
    } else {
        tile = bid % ntiles;
        split = bid / ntiles;
    }
    int ti = 0;
    while ((ti + 1) * (ti + 2) / 2 <= tile) ti++;
    const int tj = tile - ti * (ti + 1) / 2;
    const bool diag = ti == tj;
    const int rI = ti * WT, rJ = tj * WT;
    const int q0 = split * chunk;
    const int q1 = min(nact, q0 + chunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    // staging: thread t carries centroid column cl = t % 128 of the tile
    // side and the chunk's rows kb + 2u (u < 8), kb = t / 128; a wave's load
    // of one row is 64 consecutive doubles, and its LDS stores (row cl of
    // [centroid][k], stride 17) fall in distinct banks
    const int cl = threadIdx.x & 127, kb = threadIdx.x >> 7;
    double pa[8], pb[8], pw[8];
    int srcs[8];
    // source rows of this thread's chunk rows at q (-1 past the end); fetched
    // a chunk before the rows themselves, so the two loads do not chain
    auto src_of = [&](int q, int (&sr)[8]) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int qq = q + kb + 2 * u;
            sr[u] = qq < q1 ? (act ? act[qq] : qq) : -1;
        }
    };
    auto load = [&](int q, const int (&sr)[8]) {
        // raw values only: the weight is applied at the store, so nothing
        // waits for these loads before the chunk in LDS has been multiplied
        const int ca = rI + cl, cb = rJ + cl;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int qq = q + kb + 2 * u;
            const bool ok = sr[u] >= 0;
            const double *row = Phi + (size_t)(ok ? sr[u] : 0) * B;
            pw[u] = (ok && sw) ? sw[qq] : 1.0;
            pa[u] = (ok && ca < B) ? row[ca] : 0.0;
            pb[u] = (ok && !diag && cb < B) ? row[cb] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 8; u++) Sa[buf][cl][kb + 2 * u] = sw ? pa[u] * pw[u] : pa[u];
        if (!diag) {
#pragma unroll
            for (int u = 0; u < 8; u++) Sb[buf][cl][kb + 2 * u] = sw ? pb[u] * pw[u] : pb[u];
        }
    };
    v4d acc[4][4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[mb][nb] = (v4d){0.0, 0.0, 0.0, 0.0};
    src_of(q0, srcs);
    load(q0, srcs);
    src_of(q0 + WK, srcs);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int q = q0; q < q1; q += WK) {
        const bool more = q + WK < q1;
        if (more) {
            load(q + WK, srcs);
            src_of(q + 2 * WK, srcs);
        }
        double (*A)[WLP] = Sa[buf];
        double (*Bm)[WLP] = diag ? Sa[buf] : Sb[buf];
#pragma unroll
        for (int ks = 0; ks < WK / 4; ks++) {
            const int kk = ks * 4 + (lane >> 4);
            double av[4], bv[4];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) av[mb] = A[wr * 64 + mb * 16 + (lane & 15)][kk];
#pragma unroll
            for (int nb = 0; nb < 4; nb++) bv[nb] = Bm[wc * 64 + nb * 16 + (lane & 15)][kk];
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
#pragma unroll
                for (int nb = 0; nb < 4; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mb], bv[nb], acc[mb][nb], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    double *P = part + (size_t)split * B * B;
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int j = rJ + wc * 64 + nb * 16 + (lane & 15);
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = rI + wr * 64 + mb * 16 + (lane >> 4) + 4 * r;
                if (i < B && j < B && i >= j) {
                    P[(size_t)j * B + i] = acc[mb][nb][r];
                    P[(size_t)i * B + j] = acc[mb][nb][r];
                }
            }
    }
}

// K1 = KC + sum_splits part (fixed order); K1 has leading dimension ldk
__global__ void wgram_finish(int B, int nsplit, const double *__restrict__ part, const double *__restrict__ KC,
                             double *__restrict__ K1, int ldk) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t BB = (size_t)B * B;
    if (e >= BB) return;
    double s = 0.0;
    for (int p = 0; p < nsplit; p++) s += part[(size_t)p * BB + e];
    const size_t c = e / B, r = e % B;
    K1[c * ldk + r] = KC[e] + s;
}

// K2[r] = sum_q (sw_q Phi[a_q][r]) (sw_q y_q) over the active rows; one thread
// per centroid, rows in order (deterministic).  Eight rows' loads are issued
// before their eight terms are added, in row order, so each wave keeps eight
// row reads in flight.
__global__ void wvec_kernel(int B, const double *__restrict__ Phi, const int32_t *__restrict__ act,
                            const double *__restrict__ sw, const double *__restrict__ yv, double ymul, int nact,
                            int chunk, double *__restrict__ part) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int split = blockIdx.y;
    if (r >= B) return;
    const int q0 = split * chunk;
    int q1 = q0 + chunk;
    if (q1 > nact) q1 = nact;
    double s = 0.0;
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
        double x[8], w[8], yy[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int src = act ? act[q + u] : q + u;
            x[u] = Phi[(size_t)src * B + r];
            w[u] = sw ? sw[q + u] : 1.0;
            yy[u] = yv[src];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) s += (w[u] * x[u]) * (w[u] * (ymul * yy[u]));
    }
    for (; q < q1; q++) {
        const int src = act ? act[q] : q;
        const double w = sw ? sw[q] : 1.0;
        s += (w * Phi[(size_t)src * B + r]) * (w * (ymul * yv[src]));
    }
    part[(size_t)split * B + r] = s;
}

// e_i = y_i - sum_r Phi[i][r] beta_r, one wave per row
__global__ __launch_bounds__(256) void resid_rows_kernel(int n, int B, const double *__restrict__ Phi,
                                                         const double *__restrict__ beta, const double *__restrict__ yv,
                                                         double *__restrict__ e) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double *row = Phi + (size_t)i * B;
    double v = 0.0;
    for (int r = lane; r < B; r += 64) v += beta[r] * row[r];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) e[i] = yv[i] - v;
}

// D_i from e (src/budgeted-train.c:448-457) and per-block active counts
__global__ void irls_weights_kernel(int n, const double *__restrict__ e, const double *__restrict__ yv, double C,
                                    double M, double *__restrict__ D, int *__restrict__ blockcount) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double d;
        if (e[i] * yv[i] < 0.0) d = 0.0;
        else d = 1.0 * C / (yv[i] * e[i]);
        if (d > M) d = M;
        D[i] = d;
        if (d != 0.0) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) blockcount[blockIdx.x] = cnt;
}

// exclusive scan of the block counts (single work-group, sequential chunks)
__global__ void scan_counts_kernel(int nb, const int *__restrict__ bc, int *__restrict__ off, int *__restrict__ total) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = (i < nb) ? bc[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int add = (threadIdx.x >= o) ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) off[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// ordered compaction: act[off_b + rank] = i, sw = sqrt(D_i)
__global__ void compact_kernel(int n, const double *__restrict__ D, const int *__restrict__ off,
                               int32_t *__restrict__ act, double *__restrict__ sw) {
    __shared__ int flags[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = (i < n && D[i] != 0.0) ? 1 : 0;
    flags[threadIdx.x] = f;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int add = (threadIdx.x >= o) ? flags[threadIdx.x - o] : 0;
        __syncthreads();
        flags[threadIdx.x] += add;
        __syncthreads();
    }
    if (f) {
        const int pos = off[blockIdx.x] + flags[threadIdx.x] - 1;
        act[pos] = i;
        sw[pos] = sqrt(D[i]);
    }
}

}  // namespace

extern "C" {

void irwls_dev_gemm(int transA, int transB, int m, int n, int k, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    const int tiles = ((m + T - 1) / T) * ((n + T - 1) / T);
    // slices of at least 64 k until the grid has ~256 work-groups
    int nz = 1;
    while (tiles * nz * 2 <= 256 && k / (nz * 2) >= 64) nz *= 2;
    const int kper = nz > 1 ? (((k + nz - 1) / nz + BK - 1) / BK) * BK : (k > 0 ? k : 1);
    nz = (k + kper - 1) / kper > 1 ? (k + kper - 1) / kper : 1;
    double *part = nz > 1 ? (double *)scratch(1, (size_t)nz * m * n * sizeof(double)) : nullptr;
    dim3 grid((m + T - 1) / T, (n + T - 1) / T, nz);
    if (!transA && !transB) hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, dim3(256), 0, stream(), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kper, part);
    else if (transA && !transB) hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, dim3(256), 0, stream(), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kper, part);
    else if (!transA && transB) hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, dim3(256), 0, stream(), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kper, part);
    else hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, dim3(256), 0, stream(), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kper, part);
    IRWLS_LAUNCH_CHECK();
    if (nz > 1) {
        const size_t e = (size_t)m * n;
        hipLaunchKernelGGL(gemm_splitk_finish, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, stream(), m, n, nz,
                           part, alpha, beta, C, ldc);
        IRWLS_LAUNCH_CHECK();
    }
}

// Rows per partial sum of the gain kernel chosen for ncand candidates.
static int gain_tile_rows(int ncand) { return ncand <= 16 ? T : GR; }

// Launches the gain products and leaves per-row-group partial sums in
// scratch slot 10 (ncand per group, in row order); returns the rows per group
// (the row tile, or gain_split_rows with the k range split).
static int sgma_gain_parts(int n, int s, int ncand, const double *KSC, const double *Z, int ldz, const double *KSM,
                           int ldm) {
    const int tr = gain_tile_rows(ncand);
    const int rt = (n + tr - 1) / tr, ct = (ncand + T - 1) / T;
    const int rb = gain_split_rows(ncand);
    const int nbs = (n + rb - 1) / rb;
    double *part = (double *)scratch(10, (size_t)nbs * ncand * sizeof(double));
    // Few row tiles (at most 2 work-groups per CU, as on a row subsample):
    // each work-group's chain of k chunks, one dependent HBM load per chunk,
    // bounds the kernel, so the k range is split over up to 8 work-groups of
    // at least 128 k each (at most 8 work-groups per CU in all), and
    // gain_split_finish adds the ranges' products in order before the
    // squares.  With more work-groups the chunks' loads already overlap: a
    // split of the 65536-row shortlist gains (1024 work-groups) in two saved
    // less than its finish cost.
    const bool wide = ncand > 16 && (long)rt * ct >= 2 * device_cus();
    const int cg = ncand <= 16 ? 1 : wide ? ct : (ncand + 31) / 32;
    const long wgs = (long)rt * cg;
    int nz = 1;
    if (wgs <= 2L * device_cus()) {
        nz = (s + 127) / 128;
        if (nz > 8) nz = 8;
        const long cap = 8L * device_cus() / wgs;
        if (nz > cap) nz = (int)cap;
        if (nz < 1) nz = 1;
    }
    int kchunk = s;
    double *rpart = nullptr;
    if (nz > 1) {
        kchunk = ((s + nz - 1) / nz + GK - 1) / GK * GK;
        nz = (s + kchunk - 1) / kchunk;
        rpart = (double *)scratch(8, (size_t)nz * n * ncand * sizeof(double));
    }
    {
        TimerScope ts(ncand <= 16 ? IRWLS_T_GAIN16 : IRWLS_T_SGMA, 2.0 * n * (double)s * ncand,
                      8.0 * (double)n * s + 8.0 * (double)n * ncand);
        if (ncand <= 16)
            ts.launch(sgma_gain_narrow_kernel, dim3(rt, 1, nz), dim3(256), 0, n, s, ncand, KSC, Z, ldz, KSM, ldm,
                      part, kchunk, rpart);
        else if (wide)
            ts.launch(sgma_gain_kernel<T>, dim3(rt, ct, nz), dim3(256), 0, n, s, ncand, KSC, Z, ldz, KSM, ldm, part,
                      kchunk, rpart);
        else   // few row tiles: 32-candidate tiles, twice the work-groups, the same sums in the same order
            ts.launch(sgma_gain_kernel<32>, dim3(rt, cg, nz), dim3(256), 0, n, s, ncand, KSC, Z, ldz, KSM, ldm,
                      part, kchunk, rpart);
    }
    if (!rpart) return tr;
    hipLaunchKernelGGL(gain_split_finish, dim3((nbs + 3) / 4), dim3(256), 0, stream(), n, ncand, nz, rpart, KSM, ldm,
                       part);
    IRWLS_LAUNCH_CHECK();
    return rb;
}

void irwls_dev_sgma_gain(int n, int s, int ncand, const double *KSC, const double *Z, int ldz, const double *KSM,
                         int ldm, double *qout) {
    if (n <= 0 || ncand <= 0) return;
    const int g = sgma_gain_parts(n, s, ncand, KSC, Z, ldz, KSM, ldm);
    const double *part = (const double *)scratch(10, 0);
    hipLaunchKernelGGL(col_sum_kernel, dim3(ncand), dim3(256), 0, stream(), part, (n + g - 1) / g, ncand, qout);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_gain_half(int n, int s, int ncand, const double *KSC, const double *Z, int ldz, const double *KSM,
                              int ldm, double *qout, double *qhalf) {
    if (n <= 0 || ncand <= 0) return;
    const int g = sgma_gain_parts(n, s, ncand, KSC, Z, ldz, KSM, ldm);
    const double *part = (const double *)scratch(10, 0);
    const int np = (n + g - 1) / g;
    hipLaunchKernelGGL(col_sum_kernel, dim3(ncand), dim3(256), 0, stream(), part, np, ncand, qout);
    IRWLS_LAUNCH_CHECK();
    // the first half of the row tiles (at least one): the sum over their rows
    const int tr = gain_tile_rows(ncand);
    const int rt = (n + tr - 1) / tr;
    const int h = rt / 2 > 0 ? rt / 2 : 1;
    const long hp = ((long)h * tr + g - 1) / g;
    hipLaunchKernelGGL(col_sum_kernel, dim3(ncand), dim3(256), 0, stream(), part, (int)(hp < np ? hp : np), ncand,
                       qhalf);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_eta(int s, int ncand, const double *KNC, int ldk, const double *Z, int ldz, double *eta) {
    if (ncand <= 0) return;
    hipLaunchKernelGGL(sgma_eta_kernel, dim3((ncand + 3) / 4), dim3(256), 0, stream(), s, ncand, KNC, ldk, Z, ldz, eta);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_gather_col(const double *src, int ld, int col, int n, double *dst) {
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_col_kernel, dim3((n + 255) / 256), dim3(256), 0, stream(), src, ld, col, n, dst);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_scatter_row(double *M, int ld, int row, int count, const double *v) {
    if (count <= 0) return;
    hipLaunchKernelGGL(scatter_row_kernel, dim3((count + 255) / 256), dim3(256), 0, stream(), M, ld, row, count, v);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_lrow(int s, int ld, const double *kcc, const double *L2, const double *IL2, double *L,
                         double *Li) {
    if (s < 0) return;
    hipLaunchKernelGGL(sgma_lrow_kernel, dim3(1), dim3(256), 0, stream(), s, ld, kcc, L2, IL2, L, Li);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_pick(int nc, const double *q, const double *eta, const int32_t *cand, int s, int n,
                         const double *y, int32_t *cent, int32_t *st, double *value) {
    if (nc <= 0) return;
    hipLaunchKernelGGL(sgma_pick_kernel, dim3(1), dim3(64), 0, stream(), nc, q, eta, cand, s, n, y, cent, st, value);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_gather_col_at(const double *src, int ld, const int32_t *col, int n, double *dst) {
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_col_at_kernel, dim3((n + 255) / 256), dim3(256), 0, stream(), src, ld, col, n, dst);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_copy_row_at(const double *src, int ld, const int32_t *row, int count, double *dst) {
    if (count <= 0) return;
    hipLaunchKernelGGL(copy_row_at_kernel, dim3((count + 255) / 256), dim3(256), 0, stream(), src, ld, row, count,
                       dst);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_topk(int P, const float *score, int k, const int32_t *cidx, int32_t *out) {
    if (P <= 0 || k <= 0) return;
    if (P > 1024 || k > P) {
        fprintf(stderr, "libirwls: irwls_dev_sgma_topk needs k <= P <= 1024 (P=%d, k=%d)\n", P, k);
        exit(2);
    }
    hipLaunchKernelGGL(sgma_topk_kernel, dim3(1), dim3(64), 0, stream(), P, score, k, cidx, out, nullptr);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_topk_pos(int P, const float *score, int k, const int32_t *cidx, int32_t *out, int32_t *pos) {
    if (P <= 0 || k <= 0) return;
    if (P > 1024 || k > P) {
        fprintf(stderr, "libirwls: irwls_dev_sgma_topk_pos needs k <= P <= 1024 (P=%d, k=%d)\n", P, k);
        exit(2);
    }
    hipLaunchKernelGGL(sgma_topk_kernel, dim3(1), dim3(64), 0, stream(), P, score, k, cidx, out, pos);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sgma_gather_shortlist(int k, const int32_t *pos, int s, int ld, const double *KNC, const double *Z,
                                     const double *eta, double *KNC2, double *Z2, double *eta2) {
    if (k <= 0) return;
    if (s > ld) {
        fprintf(stderr, "libirwls: irwls_dev_sgma_gather_shortlist needs s <= ld (s=%d, ld=%d)\n", s, ld);
        exit(2);
    }
    hipLaunchKernelGGL(gather_shortlist_kernel, dim3(k), dim3(256), 0, stream(), s, ld, pos, KNC, Z, eta, KNC2, Z2,
                       eta2);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_wgram(int B, const double *Phi, const int32_t *act, const double *sw, int nact, const double *KC,
                     double *K1, int ldk) {
    const int tb = (B + WT - 1) / WT;
    const int ntiles = tb * (tb + 1) / 2;
    // about four rounds of two work-groups per CU (a short last round idles
    // less of the device), at least 256 rows per split, partials under 1 GiB;
    // a multiple of 8 splits when there are 8 or more
    int nsplit = (4 * 512 + ntiles - 1) / ntiles;
    if (nsplit > nact / 256) nsplit = nact / 256;
    const size_t BB = (size_t)B * B;
    while (nsplit > 1 && (size_t)nsplit * BB * sizeof(double) > ((size_t)1 << 30)) nsplit--;
    if (nsplit < 1) nsplit = 1;
    if (nsplit >= 8) {
        nsplit = (nsplit + 7) & ~7;
        while (nsplit > 8 && (size_t)nsplit * BB * sizeof(double) > ((size_t)1 << 30)) nsplit -= 8;
    }
    int chunk = (nact + nsplit - 1) / nsplit;
    chunk = (chunk + WK - 1) / WK * WK;
    if (chunk < WK) chunk = WK;
    double *part = (double *)scratch(11, (size_t)nsplit * BB * sizeof(double));
    {
        TimerScope ts(IRWLS_T_SYRKD, (double)nact * B * (double)(B + 1), 8.0 * (double)nact * B);
        if (nact > 0) {
            ts.launch(wgram_kernel, dim3(ntiles * nsplit), dim3(256), 0, B, Phi, act, sw, nact, chunk, ntiles,
                      nsplit % 8 == 0 ? 1 : 0, part);
        } else {
            irwls_dev_memset0(part, BB * sizeof(double));
        }
    }
    hipLaunchKernelGGL(wgram_finish, dim3((unsigned)((BB + 255) / 256)), dim3(256), 0, stream(), B,
                       nact > 0 ? nsplit : 1, part, KC, K1, ldk);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_wvec(int B, const double *Phi, const int32_t *act, const double *sw, const double *y, double ymul,
                    int nact, double *K2) {
    int nsplit = 64;
    int chunk = (nact + nsplit - 1) / nsplit;
    if (chunk < 1) chunk = 1;
    nsplit = nact > 0 ? (nact + chunk - 1) / chunk : 1;
    double *part = (double *)scratch(12, (size_t)nsplit * B * sizeof(double));
    if (nact > 0) {
        hipLaunchKernelGGL(wvec_kernel, dim3((B + 255) / 256, nsplit), dim3(256), 0, stream(), B, Phi, act, sw, y, ymul,
                           nact, chunk, part);
        IRWLS_LAUNCH_CHECK();
    } else {
        irwls_dev_memset0(part, (size_t)B * sizeof(double));
    }
    hipLaunchKernelGGL(col_sum_kernel, dim3(B), dim3(256), 0, stream(), part, nsplit, B, K2);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_resid_rows(int n, int B, const double *Phi, const double *beta, const double *y, double *e) {
    if (n <= 0) return;
    TimerScope ts(IRWLS_T_MATVEC, 2.0 * n * (double)B, 8.0 * (double)n * B + 16.0 * n + 8.0 * B);
    ts.launch(resid_rows_kernel, dim3((n + 3) / 4), dim3(256), 0, n, B, Phi, beta, y, e);
}

int irwls_dev_irls_weights(int n, const double *e, const double *y, double C, double M, double *D, int32_t *act,
                           double *sw) {
    if (n <= 0) return 0;
    const int nb = (n + 255) / 256;
    int *bc = (int *)scratch(13, (size_t)(2 * nb + 2) * sizeof(int));
    int *off = bc + nb;
    int *total = bc + 2 * nb;
    hipLaunchKernelGGL(irls_weights_kernel, dim3(nb), dim3(256), 0, stream(), n, e, y, C, M, D, bc);
    IRWLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, stream(), nb, bc, off, total);
    IRWLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(compact_kernel, dim3(nb), dim3(256), 0, stream(), n, D, off, act, sw);
    IRWLS_LAUNCH_CHECK();
    int h = 0;
    irwls_dev_d2h(&h, total, sizeof(int));
    return h;
}

}  // extern "C"
