// Kernel-matrix tiles for gfx950.
//
// Dense rows: k(x_i, z_j) from the norm expansion |x|^2 + |z|^2 - 2 x.z with
// the x.z block computed by FP64 MFMA (v_mfma_f64_16x16x4f64) on 64x64 tiles
// staged through LDS, and the exp / weight / row-reduction epilogue fused.
// CSR rows: lanes run over columns of a transposed dense column block, each
// row's (col, val) pairs are broadcast; same epilogues.
//
// Replaces the scalar kernelFunction loops of the reference
// (src/kernels.c:63-118; call sites src/full-train.c:178,243,336,447,504,
// src/budgeted-train.c:180,244,340,354, src/LIBIRWLS-predict.c:69,112).
#include "common.h"

using namespace irwls;

namespace {

constexpr int TM = 64;       // tile rows
constexpr int TN = 64;       // tile cols
constexpr int BK = 16;       // k chunk staged in LDS
constexpr int LDP = BK + 1;  // padded LDS row (doubles)

enum { MODE_MATVEC = 0, MODE_MATRIX = 1 };

struct DenseSide {
    const double *X;
    int ldx;
    const int32_t *idx;  // may be null
    const double *nrm;   // indexed by original row
    int n;
};

__device__ __forceinline__ double kfun(int kernel, double gamma, double dot, double nr, double nc) {
    if (kernel == 0) return dot;
    return exp(-gamma * ((nr + nc) - 2.0 * dot));
}

// Loads a TM x BK chunk of rows [r0, r0+TM) x [k0, k0+BK) into LDS (zero padded).
__device__ __forceinline__ void load_chunk(double (*S)[LDP], const DenseSide &s, int r0, int k0, int d) {
    const int t = threadIdx.x;          // 256 threads: 64 rows x 4 quads
    const int row = t >> 2;
    const int kq = (t & 3) * 4;
    const int gr = r0 + row;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (gr < s.n) {
        const int src = s.idx ? s.idx[gr] : gr;
        const double *p = s.X + (size_t)src * s.ldx + k0 + kq;
        const int kk = k0 + kq;
        if (kk + 3 < d) {
            v0 = p[0]; v1 = p[1]; v2 = p[2]; v3 = p[3];
        } else {
            if (kk < d) v0 = p[0];
            if (kk + 1 < d) v1 = p[1];
            if (kk + 2 < d) v2 = p[2];
        }
    }
    S[row][kq] = v0;
    S[row][kq + 1] = v1;
    S[row][kq + 2] = v2;
    S[row][kq + 3] = v3;
}

// One 64x64 tile of dot products, accumulated in acc[mb][nb] (v4d each):
// lane l of wave (wr, wc) owns rows wr*32 + mb*16 + l/16 + 4*r and column
// wc*32 + nb*16 + l%16 of the tile.
__device__ __forceinline__ void tile_dots(v4d acc[2][2], double (*As)[TM][LDP], double (*Bs)[TN][LDP],
                                          const DenseSide &rs, const DenseSide &cs,
                                          int r0, int c0, int d) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    for (int mb = 0; mb < 2; mb++)
        for (int nb = 0; nb < 2; nb++) acc[mb][nb] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int nchunk = (d + BK - 1) / BK;
    load_chunk(As[0], rs, r0, 0, d);
    load_chunk(Bs[0], cs, c0, 0, d);
    __syncthreads();
    for (int ch = 0; ch < nchunk; ch++) {
        const int cur = ch & 1;
        if (ch + 1 < nchunk) {
            load_chunk(As[cur ^ 1], rs, r0, (ch + 1) * BK, d);
            load_chunk(Bs[cur ^ 1], cs, c0, (ch + 1) * BK, d);
        }
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            const int kk = ks * 4 + (lane >> 4);
            double a0 = As[cur][wr * 32 + (lane & 15)][kk];
            double a1 = As[cur][wr * 32 + 16 + (lane & 15)][kk];
            double b0 = Bs[cur][wc * 32 + (lane & 15)][kk];
            double b1 = Bs[cur][wc * 32 + 16 + (lane & 15)][kk];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double row_nrm(const DenseSide &s, int i) {
    return s.nrm ? s.nrm[s.idx ? s.idx[i] : i] : 0.0;
}

// Fused kernel-block x vector.  grid = (row tiles, column splits).  Split s
// covers column tiles [s*tps, (s+1)*tps).  Writes partial sums part[s*n + i]
// when nsplit > 1, else applies the output epilogue directly.
__global__ __launch_bounds__(256) void kmv_dense_kernel(
    DenseSide rs, DenseSide cs, int d, int kernel, double gamma, const double *__restrict__ w,
    int tps, int nsplit, double *__restrict__ part,
    double a, double b, double cst, const double *__restrict__ rsc, double *__restrict__ out) {
    __shared__ double As[2][TM][LDP];
    __shared__ double Bs[2][TN][LDP];
    __shared__ double red[2][TM];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = blockIdx.x * TM;
    const int ct_begin = blockIdx.y * tps;
    int ct_end = ct_begin + tps;
    const int ntiles = (cs.n + TN - 1) / TN;
    if (ct_end > ntiles) ct_end = ntiles;

    // per-lane row partial sums: rows wr*32 + mb*16 + lane/16 + 4*r
    double rsum[2][4];
    double nr[2][4];
    for (int mb = 0; mb < 2; mb++)
        for (int r = 0; r < 4; r++) {
            rsum[mb][r] = 0.0;
            const int gi = r0 + wr * 32 + mb * 16 + (lane >> 4) + 4 * r;
            nr[mb][r] = (gi < rs.n) ? row_nrm(rs, gi) : 0.0;
        }
    for (int ct = ct_begin; ct < ct_end; ct++) {
        const int c0 = ct * TN;
        v4d acc[2][2];
        tile_dots(acc, As, Bs, rs, cs, r0, c0, d);
        for (int nb = 0; nb < 2; nb++) {
            const int gj = c0 + wc * 32 + nb * 16 + (lane & 15);
            if (gj < cs.n) {
                const double wj = w[gj];
                const double ncj = row_nrm(cs, gj);
                for (int mb = 0; mb < 2; mb++)
                    for (int r = 0; r < 4; r++)
                        rsum[mb][r] += wj * kfun(kernel, gamma, acc[mb][nb][r], nr[mb][r], ncj);
            }
        }
    }
    // reduce over the 16 lanes that share rows (lane % 16 differs)
    for (int mb = 0; mb < 2; mb++)
        for (int r = 0; r < 4; r++) {
            double v = rsum[mb][r];
            v += __shfl_xor(v, 1);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 4);
            v += __shfl_xor(v, 8);
            rsum[mb][r] = v;
        }
    if ((lane & 15) == 0) {
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) red[wc][wr * 32 + mb * 16 + (lane >> 4) + 4 * r] = rsum[mb][r];
    }
    __syncthreads();
    if (threadIdx.x < TM) {
        const int i = threadIdx.x;
        const int gi = r0 + i;
        if (gi < rs.n) {
            const double s = red[0][i] + red[1][i];
            if (nsplit > 1) {
                part[(size_t)blockIdx.y * rs.n + gi] = s;
            } else {
                const double sc = rsc ? rsc[gi] : 1.0;
                out[gi] = (a == 0.0 ? 0.0 : a * out[gi]) + sc * (b * s) + cst;
            }
        }
    }
}

// Materialised kernel block K[i*ldk + j].
__global__ __launch_bounds__(256) void kmat_dense_kernel(DenseSide rs, DenseSide cs, int d, int kernel,
                                                         double gamma, double *__restrict__ K, int ldk) {
    __shared__ double As[2][TM][LDP];
    __shared__ double Bs[2][TN][LDP];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = blockIdx.x * TM;
    const int c0 = blockIdx.y * TN;
    v4d acc[2][2];
    tile_dots(acc, As, Bs, rs, cs, r0, c0, d);
    for (int nb = 0; nb < 2; nb++) {
        const int gj = c0 + wc * 32 + nb * 16 + (lane & 15);
        if (gj >= cs.n) continue;
        const double ncj = row_nrm(cs, gj);
        for (int mb = 0; mb < 2; mb++)
            for (int r = 0; r < 4; r++) {
                const int gi = r0 + wr * 32 + mb * 16 + (lane >> 4) + 4 * r;
                if (gi < rs.n)
                    K[(size_t)gi * ldk + gj] = kfun(kernel, gamma, acc[mb][nb][r], row_nrm(rs, gi), ncj);
            }
    }
}

// Materialised block for at most 16 columns (a shortlist of candidates):
// one thread per CSR row with every column's dot product in registers, so a
// narrow block costs one pass over the rows' nonzeros instead of one wave per
// row with most lanes idle.  A nonzero whose feature no column has (mask 0)
// adds exact zeros and is skipped, as in csr_row_dot.
constexpr int NARROW = 16;

__global__ __launch_bounds__(256) void kmat_csr_narrow_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rnrm, const int32_t *__restrict__ ridx, int nrows,
    const double *__restrict__ XcT, int ldc, int d, const unsigned long long *__restrict__ cmask,
    const double *__restrict__ cnrm, int ncols, int kernel, double gamma, double *__restrict__ K, int ldk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int src = ridx ? ridx[i] : i;
    const int64_t p0 = rowptr[src], p1 = rowptr[src + 1];
    double acc[NARROW];
#pragma unroll
    for (int c = 0; c < NARROW; c++) acc[c] = 0.0;
    // four nonzeros per step: their columns, values and masks load together
    for (int64_t p = p0; p < p1; p += 4) {
        int f[4];
        double v[4];
        unsigned m[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool ok = p + u < p1;
            f[u] = ok ? col[p + u] : d;
            v[u] = ok ? val[p + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) m[u] = f[u] < d ? (unsigned)cmask[f[u]] : 0u;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (m[u] == 0u) continue;
            const double *xr = XcT + (size_t)f[u] * ldc;
#pragma unroll
            for (int c = 0; c < NARROW; c++)
                if ((m[u] >> c) & 1u) acc[c] = fma(v[u], xr[c], acc[c]);
        }
    }
    const double nr = rnrm ? rnrm[src] : 0.0;
#pragma unroll
    for (int c = 0; c < NARROW; c++)
        if (c < ncols) K[(size_t)i * ldk + c] = kfun(kernel, gamma, acc[c], nr, cnrm ? cnrm[c] : 0.0);
}

// The same product for a block of 256 consecutive rows (no row selection):
// the block's nonzeros are one contiguous range, brought into LDS in chunks
// of BCH with coalesced loads, and each thread walks its row's part of each
// chunk in order.  The per-thread walk of the kernel above reads its row
// straight from memory, so every wave-wide load touches 64 cache lines for
// 12 useful bytes each; here the nonzeros cross HBM and L2 once.
constexpr int BCH = 2048;

__global__ __launch_bounds__(256) void kmat_csr_block_narrow_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rnrm, int nrows, const double *__restrict__ XcT, int ldc, int d,
    const unsigned long long *__restrict__ cmask, const double *__restrict__ cnrm, int ncols, int kernel,
    double gamma, double *__restrict__ K, int ldk) {
    __shared__ int s_col[BCH];
    __shared__ double s_val[BCH];
    const int r0 = blockIdx.x * 256;
    const int i = r0 + threadIdx.x;
    const int rend = min(nrows, r0 + 256);
    const int64_t P0 = rowptr[r0], P1 = rowptr[rend];
    const bool live = i < nrows;
    const int64_t p0 = live ? rowptr[i] : 0, p1 = live ? rowptr[i + 1] : 0;
    double acc[NARROW];
#pragma unroll
    for (int c = 0; c < NARROW; c++) acc[c] = 0.0;
    for (int64_t cb = P0; cb < P1; cb += BCH) {
        const int clen = (int)min((int64_t)BCH, P1 - cb);
        __syncthreads();   // the previous chunk has been read
        {
            // all of this thread's loads in flight before the first store
            int cv[BCH / 256];
            double vv[BCH / 256];
#pragma unroll
            for (int u = 0; u < BCH / 256; u++) {
                const int k = threadIdx.x + 256 * u;
                cv[u] = k < clen ? col[cb + k] : 0;
                vv[u] = k < clen ? val[cb + k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < BCH / 256; u++) {
                const int k = threadIdx.x + 256 * u;
                if (k < clen) {
                    s_col[k] = cv[u];
                    s_val[k] = vv[u];
                }
            }
        }
        __syncthreads();
        const int64_t b = max(p0, cb), e = min(p1, cb + clen);
        // four nonzeros per step, their masks fetched together
        for (int64_t p = b; p < e; p += 4) {
            int f[4];
            double v[4];
            unsigned m[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const bool ok = p + u < e;
                f[u] = ok ? s_col[p + u - cb] : d;
                v[u] = ok ? s_val[p + u - cb] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) m[u] = f[u] < d ? (unsigned)cmask[f[u]] : 0u;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (m[u] == 0u) continue;
                const double *xr = XcT + (size_t)f[u] * ldc;
#pragma unroll
                for (int c = 0; c < NARROW; c++)
                    if ((m[u] >> c) & 1u) acc[c] = fma(v[u], xr[c], acc[c]);
            }
        }
    }
    if (!live) return;
    const double nr = rnrm ? rnrm[i] : 0.0;
#pragma unroll
    for (int c = 0; c < NARROW; c++)
        if (c < ncols) K[(size_t)i * ldk + c] = kfun(kernel, gamma, acc[c], nr, cnrm ? cnrm[c] : 0.0);
}

// Deterministic split reduction + output epilogue.
__global__ void kmv_finish_kernel(const double *__restrict__ part, int n, int nsplit, double a, double b,
                                  double cst, const double *__restrict__ rsc, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; k++) s += part[(size_t)k * n + i];
    const double sc = rsc ? rsc[i] : 1.0;
    out[i] = (a == 0.0 ? 0.0 : a * out[i]) + sc * (b * s) + cst;
}

// ---- CSR rows x transposed dense columns ---------------------------------
// XcT is d x ldc; cmask[(j / 64) * d + f] has bit j % 64 set when column j's
// feature f is nonzero.  One wave per row, lanes over 64 columns per chunk.
// The row's nonzeros and their features' masks are fetched 64 at a time with
// one coalesced load each; the nonzeros whose feature some column of the
// chunk has (a ballot over the masks) are then visited in order, four at a
// time: the lanes whose column has the feature load their XcT entries (four
// independent loads in flight), the others take 0, and the four products
// are added in nonzero order, one fused multiply-add each.  The skipped
// products are exact zeros, so the sum is that of the full dense-column dot
// product, bit for bit.
__device__ __forceinline__ double csr_row_dot(const int32_t *__restrict__ col, const double *__restrict__ val,
                                              int64_t p0, int64_t p1, const double *__restrict__ XcT, int ldc,
                                              int d, const unsigned long long *__restrict__ cmask, int j) {
    const int lane = threadIdx.x & 63;
    double dot = 0.0;
    for (int64_t b = p0; b < p1; b += 64) {
        const int cnt = (int)min((int64_t)64, p1 - b);
        int fl = 0;
        double vl = 0.0;
        unsigned long long ml = 0;
        if (lane < cnt) {
            fl = col[b + lane];
            vl = val[b + lane];
            if (fl < d) ml = cmask[fl];
        }
        unsigned long long act = __ballot(ml != 0ull);
        while (act) {
            int qs[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (act) {
                    qs[u] = __builtin_ctzll(act);
                    act &= act - 1;
                } else {
                    qs[u] = -1;
                }
            }
            double x[4], v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                x[u] = 0.0;
                v[u] = 0.0;
                if (qs[u] >= 0) {
                    const int q = qs[u];
                    const unsigned lo = __builtin_amdgcn_readlane((unsigned)ml, q);
                    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(ml >> 32), q);
                    const unsigned bit = lane < 32 ? (lo >> lane) : (hi >> (lane - 32));
                    const int f = __builtin_amdgcn_readlane(fl, q);
                    v[u] = rdl(vl, q);
                    if (bit & 1u) x[u] = XcT[(size_t)f * ldc + j];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (qs[u] >= 0) dot = fma(v[u], x[u], dot);
        }
    }
    return dot;
}

__global__ __launch_bounds__(256) void kmv_csr_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rnrm, const int32_t *__restrict__ ridx, int nrows,
    const double *__restrict__ XcT, int ldc, int d, const unsigned long long *__restrict__ cmask,
    const double *__restrict__ cnrm, int ncols, int kernel, double gamma, const double *__restrict__ w,
    double a, double b, double cst, const double *__restrict__ rsc, double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nrows) return;
    const int src = ridx ? ridx[i] : i;
    const int64_t p0 = rowptr[src], p1 = rowptr[src + 1];
    const double nr = rnrm ? rnrm[src] : 0.0;
    double acc = 0.0;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        const int j = c0 + lane;
        const bool jok = j < ncols;
        const double dot = csr_row_dot(col, val, p0, p1, XcT, ldc, d, cmask + (size_t)(c0 >> 6) * d, j);
        if (jok) acc += w[j] * kfun(kernel, gamma, dot, nr, cnrm ? cnrm[j] : 0.0);
    }
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        const double sc = rsc ? rsc[i] : 1.0;
        out[i] = (a == 0.0 ? 0.0 : a * out[i]) + sc * (b * acc) + cst;
    }
}

__global__ __launch_bounds__(256) void kmat_csr_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rnrm, const int32_t *__restrict__ ridx, int nrows,
    const double *__restrict__ XcT, int ldc, int d, const unsigned long long *__restrict__ cmask,
    const double *__restrict__ cnrm, int ncols, int kernel, double gamma, double *__restrict__ K, int ldk) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nrows) return;
    const int src = ridx ? ridx[i] : i;
    const int64_t p0 = rowptr[src], p1 = rowptr[src + 1];
    const double nr = rnrm ? rnrm[src] : 0.0;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        const int j = c0 + lane;
        const bool jok = j < ncols;
        const double dot = csr_row_dot(col, val, p0, p1, XcT, ldc, d, cmask + (size_t)(c0 >> 6) * d, j);
        if (jok) K[(size_t)i * ldk + j] = kfun(kernel, gamma, dot, nr, cnrm ? cnrm[j] : 0.0);
    }
}

// Column block as a transposed dense matrix XcT (d x nc) + compact norms.
__global__ void cols_T_from_dense(const double *__restrict__ X, int ldx, const int32_t *__restrict__ idx,
                                  const double *__restrict__ nrm, int nc, int d, double *__restrict__ XcT,
                                  double *__restrict__ cn, unsigned long long *__restrict__ mask) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (j >= nc) return;
    const int src = idx ? idx[j] : j;
    const double x = X[(size_t)src * ldx + f];
    XcT[(size_t)f * nc + j] = x;
    if (x != 0.0) atomicOr(&mask[(size_t)(j >> 6) * d + f], 1ull << (j & 63));
    if (f == 0) cn[j] = nrm ? nrm[src] : 0.0;
}

// Scatter CSR rows into a zero-filled dense block: transposed (d x nc) when
// trans != 0, else row-major (nc x d); with mask, set the feature masks of
// the transposed block (bit j % 64 of mask[(j / 64) * d + f]).
__global__ void cols_from_csr(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                              const double *__restrict__ val, const int32_t *__restrict__ idx,
                              const double *__restrict__ nrm, int nc, int d, int trans,
                              double *__restrict__ out, double *__restrict__ cn,
                              unsigned long long *__restrict__ mask) {
    const int j = blockIdx.x;
    if (j >= nc) return;
    const int src = idx ? idx[j] : j;
    const int64_t p0 = rowptr[src], p1 = rowptr[src + 1];
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const int f = col[p];
        if (f < d) {
            if (trans) {
                out[(size_t)f * nc + j] = val[p];
                if (mask) atomicOr(&mask[(size_t)(j >> 6) * d + f], 1ull << (j & 63));
            }
            else out[(size_t)j * d + f] = val[p];
        }
    }
    if (threadIdx.x == 0) cn[j] = nrm ? nrm[src] : 0.0;
}

__global__ void row_norms_kernel(const double *__restrict__ X, int ldx, int n, int d, double *__restrict__ nrm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *p = X + (size_t)i * ldx;
    // separately rounded multiply and add (no FMA contraction), the order and
    // rounding of the readers' running sum of squares
    double s = 0.0;
    for (int k = 0; k < d; k++) s = __dadd_rn(s, __dmul_rn(p[k], p[k]));
    nrm[i] = s;
}

__global__ void axpb_kernel(int n, double a, double cst, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (a == 0.0 ? 0.0 : a * out[i]) + cst;
}

DenseSide dense_side(const irwls_rows *r) {
    DenseSide s;
    s.X = r->X;
    s.ldx = r->ldx;
    s.idx = r->idx;
    s.nrm = r->nrm;
    s.n = r->n;
    return s;
}

// Columns given as CSR rows are scattered into a dense row-major block so the
// MFMA path can use them.  Returns a DenseSide over scratch memory.
DenseSide dense_cols(const irwls_rows *c, int d, int slot_vals, int slot_nrm) {
    if (c->X) return dense_side(c);
    double *buf = (double *)scratch(slot_vals, (size_t)c->n * d * sizeof(double) + 8);
    double *cn = (double *)scratch(slot_nrm, (size_t)c->n * sizeof(double) + 8);
    irwls_dev_memset0(buf, (size_t)c->n * d * sizeof(double));
    if (c->n > 0) {
        hipLaunchKernelGGL(cols_from_csr, dim3(c->n), dim3(64), 0, stream(), c->rowptr, c->col, c->val,
                           c->idx, c->nrm, c->n, d, 0, buf, cn, nullptr);
        IRWLS_LAUNCH_CHECK();
    }
    DenseSide s;
    s.X = buf;
    s.ldx = d;
    s.idx = nullptr;
    s.nrm = cn;
    s.n = c->n;
    return s;
}

// Transposed dense column block for the CSR row kernels, with the feature
// masks of its 64-column chunks.
void cols_T(const irwls_rows *c, int d, double **XcT, double **cn, unsigned long long **mask) {
    const size_t nchunk = (size_t)(c->n + 63) / 64;
    *XcT = (double *)scratch(4, (size_t)c->n * d * sizeof(double) + 8);
    *cn = (double *)scratch(5, (size_t)c->n * sizeof(double) + 8);
    *mask = (unsigned long long *)scratch(15, nchunk * d * sizeof(unsigned long long) + 8);
    if (c->n == 0) return;
    irwls_dev_memset0(*mask, nchunk * d * sizeof(unsigned long long));
    if (c->X) {
        hipLaunchKernelGGL(cols_T_from_dense, dim3((c->n + 255) / 256, d), dim3(256), 0, stream(), c->X,
                           c->ldx, c->idx, c->nrm, c->n, d, *XcT, *cn, *mask);
    } else {
        irwls_dev_memset0(*XcT, (size_t)c->n * d * sizeof(double));
        hipLaunchKernelGGL(cols_from_csr, dim3(c->n), dim3(64), 0, stream(), c->rowptr, c->col, c->val,
                           c->idx, c->nrm, c->n, d, 1, *XcT, *cn, *mask);
    }
    IRWLS_LAUNCH_CHECK();
}

}  // namespace

extern "C" {

void irwls_dev_row_norms(const double *X, int ldx, int n, int d, double *nrm) {
    if (n <= 0) return;
    hipLaunchKernelGGL(row_norms_kernel, dim3((n + 255) / 256), dim3(256), 0, stream(), X, ldx, n, d, nrm);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_kernel_matvec(const irwls_rows *r, const irwls_rows *c, int d, int kernel, double gamma,
                             const double *w, double a, double b, double cst, const double *rs, double *out) {
    if (r->n <= 0) return;
    if (c->n <= 0) {
        hipLaunchKernelGGL(axpb_kernel, dim3((r->n + 255) / 256), dim3(256), 0, stream(), r->n, a, cst, out);
        IRWLS_LAUNCH_CHECK();
        return;
    }
    const double flops = 2.0 * r->n * (double)c->n * d;
    if (r->X) {
        DenseSide rsd = dense_side(r);
        DenseSide csd = dense_cols(c, d, 6, 7);
        const int rt = (r->n + TM - 1) / TM;
        const int ct = (c->n + TN - 1) / TN;
        // enough work groups to fill 256 CUs a few times over
        int nsplit = 1;
        while (rt * nsplit < 1024 && nsplit * 2 <= ct) nsplit *= 2;
        const int tps = (ct + nsplit - 1) / nsplit;
        nsplit = (ct + tps - 1) / tps;
        double *part = nullptr;
        if (nsplit > 1) part = (double *)scratch(8, (size_t)nsplit * r->n * sizeof(double));
        const double bytes = 8.0 * ((double)r->n * d + (double)c->n * d + 2.0 * r->n);
        TimerScope ts(IRWLS_T_KMV, flops, bytes);
        hipLaunchKernelGGL(kmv_dense_kernel, dim3(rt, nsplit), dim3(256), 0, stream(), rsd, csd, d, kernel,
                           gamma, w, tps, nsplit, part, a, b, cst, rs, out);
        IRWLS_LAUNCH_CHECK();
        if (nsplit > 1) {
            hipLaunchKernelGGL(kmv_finish_kernel, dim3((r->n + 255) / 256), dim3(256), 0, stream(), part, r->n,
                               nsplit, a, b, cst, rs, out);
            IRWLS_LAUNCH_CHECK();
        }
    } else {
        double *XcT, *cn;
        unsigned long long *cm;
        cols_T(c, d, &XcT, &cn, &cm);
        // algorithmic: the rows' nonzeros (value + column), their row pointers, the column table, the output
        const double nz = r->nnz_per_row * r->n;
        TimerScope ts(IRWLS_T_CSR, 2.0 * nz * c->n, 12.0 * nz + 8.0 * c->n * (double)d + 24.0 * r->n);
        hipLaunchKernelGGL(kmv_csr_kernel, dim3((r->n + 3) / 4), dim3(256), 0, stream(), r->rowptr, r->col, r->val,
                           r->nrm, r->idx, r->n, XcT, c->n, d, cm, cn, c->n, kernel, gamma, w, a, b, cst, rs, out);
        IRWLS_LAUNCH_CHECK();
    }
}

void irwls_dev_kernel_matrix(const irwls_rows *r, const irwls_rows *c, int d, int kernel, double gamma,
                             double *K, int ldk) {
    if (r->n <= 0 || c->n <= 0) return;
    const double flops = 2.0 * r->n * (double)c->n * d;
    if (r->X) {
        DenseSide rsd = dense_side(r);
        DenseSide csd = dense_cols(c, d, 6, 7);
        const double bytes = 8.0 * ((double)r->n * d + (double)c->n * d + (double)r->n * c->n);
        TimerScope ts(IRWLS_T_KMAT, flops, bytes);
        hipLaunchKernelGGL(kmat_dense_kernel, dim3((r->n + TM - 1) / TM, (c->n + TN - 1) / TN), dim3(256), 0,
                           stream(), rsd, csd, d, kernel, gamma, K, ldk);
        IRWLS_LAUNCH_CHECK();
    } else if (c->n <= NARROW) {
        double *XcT, *cn;
        unsigned long long *cm;
        cols_T(c, d, &XcT, &cn, &cm);
        const double nz = r->nnz_per_row * r->n;
        TimerScope ts(IRWLS_T_CSR, 2.0 * nz * c->n, 12.0 * nz + 8.0 * c->n * (double)d + 8.0 * r->n * (2.0 + c->n));
        if (r->idx)
            hipLaunchKernelGGL(kmat_csr_narrow_kernel, dim3((r->n + 255) / 256), dim3(256), 0, stream(), r->rowptr,
                               r->col, r->val, r->nrm, r->idx, r->n, XcT, c->n, d, cm, cn, c->n, kernel, gamma, K,
                               ldk);
        else
            hipLaunchKernelGGL(kmat_csr_block_narrow_kernel, dim3((r->n + 255) / 256), dim3(256), 0, stream(),
                               r->rowptr, r->col, r->val, r->nrm, r->n, XcT, c->n, d, cm, cn, c->n, kernel, gamma,
                               K, ldk);
        IRWLS_LAUNCH_CHECK();
    } else {
        double *XcT, *cn;
        unsigned long long *cm;
        cols_T(c, d, &XcT, &cn, &cm);
        const double nz = r->nnz_per_row * r->n;
        TimerScope ts(IRWLS_T_CSR, 2.0 * nz * c->n, 12.0 * nz + 8.0 * c->n * (double)d + 8.0 * r->n * (2.0 + c->n));
        hipLaunchKernelGGL(kmat_csr_kernel, dim3((r->n + 3) / 4), dim3(256), 0, stream(), r->rowptr, r->col, r->val,
                           r->nrm, r->idx, r->n, XcT, c->n, d, cm, cn, c->n, kernel, gamma, K, ldk);
        IRWLS_LAUNCH_CHECK();
    }
}

}  // extern "C"
