// Kernel-matrix tiles for gfx950.
//
// Dense rows: k(x_i, z_j) from the norm expansion |x|^2 + |z|^2 - 2 x.z with
// the x.z block computed by FP64 MFMA (v_mfma_f64_16x16x4f64) on 128x128
// tiles staged through LDS with the next k chunk's loads in flight, and the
// exp / weight / row-reduction epilogue fused.
// CSR rows: lanes run over columns of a transposed dense column block, each
// row's (col, val) pairs are broadcast; same epilogues.
//
// Replaces the scalar kernelFunction loops of the reference
// (src/kernels.c:63-118; call sites src/full-train.c:178,243,336,447,504,
// src/budgeted-train.c:180,244,340,354, src/LIBIRWLS-predict.c:69,112).
#include "common.h"

using namespace irwls;

namespace {

constexpr int TB = 128;      // tile rows / columns per work-group
constexpr int BK = 16;       // k chunk staged in LDS
constexpr int LDP = BK + 1;  // padded LDS row (doubles)

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

// One k chunk of a 128-row tile in registers: thread t holds row t/2,
// features k0 + 8 (t%2) .. +7 (zero beyond d and past the last row).
__device__ __forceinline__ void load_rows(double (&v)[8], const DenseSide &s, int r0, int k0, int d) {
    const int row = threadIdx.x >> 1, kh = k0 + (threadIdx.x & 1) * 8, g = r0 + row;
    if (g < s.n && kh < d) {
        const double *p = s.X + (size_t)(s.idx ? s.idx[g] : g) * s.ldx + kh;
        if (kh + 8 <= d) {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = p[q];
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = (kh + q < d) ? p[q] : 0.0;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = 0.0;
    }
}

__device__ __forceinline__ void store_rows(double (*S)[LDP], const double (&v)[8]) {
    const int row = threadIdx.x >> 1, kh = (threadIdx.x & 1) * 8;
#pragma unroll
    for (int q = 0; q < 8; q++) S[row][kh + q] = v[q];
}

// One 128 x 128 tile of dot products on FP64 MFMA: acc[mb][nb][r] = a_i . b_j
// for i = a0 + 64 wa + 16 mb + l/16 + 4 r of side A and j = b0 + 64 wb +
// 16 nb + l%16 of side B (wave 2 wa + wb, lane l; the accumulator layout of
// v_mfma_f64_16x16x4f64).  k runs in chunks of 16 through double-buffered LDS;
// the next chunk's global loads are issued before the current chunk's MFMAs
// and written to LDS after them, so their latency is hidden.
__device__ __forceinline__ void tile_dots(v4d (&acc)[4][4], double (*As)[TB][LDP], double (*Bs)[TB][LDP],
                                          const DenseSide &sa, const DenseSide &sb, int a0, int b0, int d) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wa = wave >> 1, wb = wave & 1;
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[mb][nb] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int nchunk = (d + BK - 1) / BK;
    double va[8], vb[8];
    load_rows(va, sa, a0, 0, d);
    load_rows(vb, sb, b0, 0, d);
    __syncthreads();   // the previous tile's last chunk has been read
    store_rows(As[0], va);
    store_rows(Bs[0], vb);
    __syncthreads();
    for (int ch = 0; ch < nchunk; ch++) {
        const int cur = ch & 1;
        const bool more = ch + 1 < nchunk;
        if (more) {
            load_rows(va, sa, a0, (ch + 1) * BK, d);
            load_rows(vb, sb, b0, (ch + 1) * BK, d);
        }
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            const int kk = ks * 4 + (lane >> 4);
            double fa[4], fb[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                fa[q] = As[cur][wa * 64 + q * 16 + (lane & 15)][kk];
                fb[q] = Bs[cur][wb * 64 + q * 16 + (lane & 15)][kk];
            }
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
#pragma unroll
                for (int nb = 0; nb < 4; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[mb], fb[nb], acc[mb][nb], 0, 0, 0);
        }
        if (more) {
            // the other buffer was last read in the previous chunk, before its barrier
            store_rows(As[cur ^ 1], va);
            store_rows(Bs[cur ^ 1], vb);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double row_nrm(const DenseSide &s, int i) {
    return s.nrm ? s.nrm[s.idx ? s.idx[i] : i] : 0.0;
}

// Fused kernel-block x vector.  grid = (row tiles, column splits); split s
// covers column tiles [s*tps, (s+1)*tps).  The columns are the MFMA A side,
// so each lane keeps four fixed rows (r0 + 64 wb + 16 nb + l%16) and sums
// w_j k_ij over its columns in registers; the lanes sharing a row are
// reduced once at the end.  Writes partial sums part[s*n + i] when
// nsplit > 1, else applies the output epilogue directly.
__global__ __launch_bounds__(256, 2) void kmv_dense_kernel(
    DenseSide rs, DenseSide cs, int d, int kernel, double gamma, const double *__restrict__ w,
    int tps, int nsplit, double *__restrict__ part,
    double a, double b, double cst, const double *__restrict__ rsc, double *__restrict__ out) {
    __shared__ double As[2][TB][LDP];
    __shared__ double Bs[2][TB][LDP];
    __shared__ __align__(16) double wn[2][2][TB];   // [tile parity][w, column norm][permuted column]
    __shared__ double red[2][TB];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wa = wave >> 1, wb = wave & 1;
    const int r0 = blockIdx.x * TB;
    const int ct_begin = blockIdx.y * tps;
    const int ntiles = (cs.n + TB - 1) / TB;
    const int ct_end = ct_begin + tps < ntiles ? ct_begin + tps : ntiles;
    double nr[4], rsum[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int gi = r0 + wb * 64 + nb * 16 + (lane & 15);
        nr[nb] = gi < rs.n ? row_nrm(rs, gi) : 0.0;
        rsum[nb] = 0.0;
    }
    for (int ct = ct_begin; ct < ct_end; ct++) {
        const int c0 = ct * TB, par = (ct - ct_begin) & 1;
        if (threadIdx.x < TB) {
            // column j of the tile is kept at 16 (j/16) + 4 (j%4) + (j%16)/4, so the lane with l/16 = q
            // finds its columns q + 4 r (r = 0..3) of each 16-column block as one contiguous run; the
            // parity buffer written here was last read two tiles ago, before this tile's barriers
            const int j = threadIdx.x, gj = c0 + j;
            const int pos = (j & ~15) + 4 * (j & 3) + ((j & 15) >> 2);
            wn[par][0][pos] = gj < cs.n ? w[gj] : 0.0;
            wn[par][1][pos] = gj < cs.n ? row_nrm(cs, gj) : 0.0;
        }
        v4d acc[4][4];
        tile_dots(acc, As, Bs, cs, rs, c0, r0, d);
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const int p0 = wa * 64 + mb * 16 + 4 * (lane >> 4);
            const double2 w01 = *reinterpret_cast<const double2 *>(&wn[par][0][p0]);
            const double2 w23 = *reinterpret_cast<const double2 *>(&wn[par][0][p0 + 2]);
            const double2 n01 = *reinterpret_cast<const double2 *>(&wn[par][1][p0]);
            const double2 n23 = *reinterpret_cast<const double2 *>(&wn[par][1][p0 + 2]);
            const double wv[4] = {w01.x, w01.y, w23.x, w23.y};
            const double nv[4] = {n01.x, n01.y, n23.x, n23.y};
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
#pragma unroll
                for (int r = 0; r < 4; r++) rsum[nb] += wv[r] * kfun(kernel, gamma, acc[mb][nb][r], nr[nb], nv[r]);
        }
    }
    // the lanes l, l^16, l^32, l^48 hold the same rows
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        double v = rsum[nb];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        rsum[nb] = v;
    }
    if (lane < 16) {
#pragma unroll
        for (int nb = 0; nb < 4; nb++) red[wa][wb * 64 + nb * 16 + lane] = rsum[nb];
    }
    __syncthreads();
    if (threadIdx.x < TB) {
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

// Materialised kernel block K[i*ldk + j]; rows on the MFMA A side, so each
// register of a lane stores 16 consecutive columns across the lanes.
__global__ __launch_bounds__(256, 2) void kmat_dense_kernel(DenseSide rs, DenseSide cs, int d, int kernel,
                                                            double gamma, double *__restrict__ K, int ldk) {
    __shared__ double As[2][TB][LDP];
    __shared__ double Bs[2][TB][LDP];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wa = wave >> 1, wb = wave & 1;
    const int r0 = blockIdx.x * TB;
    const int c0 = blockIdx.y * TB;
    v4d acc[4][4];
    tile_dots(acc, As, Bs, rs, cs, r0, c0, d);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int gj = c0 + wb * 64 + nb * 16 + (lane & 15);
        if (gj >= cs.n) continue;
        const double ncj = row_nrm(cs, gj);
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gi = r0 + wa * 64 + mb * 16 + (lane >> 4) + 4 * r;
                if (gi < rs.n)
                    K[(size_t)gi * ldk + gj] = kfun(kernel, gamma, acc[mb][nb][r], row_nrm(rs, gi), ncj);
            }
    }
}

// Materialised block for at most 16 columns (a shortlist of candidates) on
// selected rows: a group of G lanes per row (G = the column count rounded up
// to a power of two), lane c of the group column c.  Each lane walks the
// row's nonzeros in order and adds v x_c where the column has the feature,
// the same fused multiply-adds in the same order as one thread per row with
// every column in registers; the group's lanes load the same nonzero (one
// transaction) and adjacent column-table entries.  A nonzero whose feature
// no column has (mask 0) adds exact zeros and is skipped, as in csr_row_dot.
constexpr int NARROW = 16;

__global__ __launch_bounds__(256) void kmat_csr_narrow_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ rnrm, const int32_t *__restrict__ ridx, int nrows,
    const double *__restrict__ XcT, int ldc, int d, const unsigned long long *__restrict__ cmask,
    const double *__restrict__ cnrm, int ncols, int glog, int kernel, double gamma, double *__restrict__ K,
    int ldk) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t >> glog, c = t & ((1 << glog) - 1);
    if (i >= nrows) return;
    const int src = ridx ? ridx[i] : i;
    const int64_t p0 = rowptr[src], p1 = rowptr[src + 1];
    double acc = 0.0;
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
        for (int u = 0; u < 4; u++) m[u] = (unsigned)f[u] < (unsigned)d ? (unsigned)cmask[f[u]] : 0u;
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) x[u] = ((m[u] >> c) & 1u) ? XcT[(size_t)f[u] * ldc + c] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if ((m[u] >> c) & 1u) acc = fma(v[u], x[u], acc);
    }
    if (c < ncols) {
        const double nr = rnrm ? rnrm[src] : 0.0;
        K[(size_t)i * ldk + c] = kfun(kernel, gamma, acc, nr, cnrm ? cnrm[c] : 0.0);
    }
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
            for (int u = 0; u < 4; u++) m[u] = (unsigned)f[u] < (unsigned)d ? (unsigned)cmask[f[u]] : 0u;
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
            if ((unsigned)fl < (unsigned)d) ml = cmask[fl];
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
    // grid.y strides over the 64-column chunks (one chunk per wave when the
    // grid has a y extent per chunk)
    for (int c0 = blockIdx.y * 64; c0 < ncols; c0 += gridDim.y * 64) {
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
        if ((unsigned)f < (unsigned)d) {
            if (trans) {
                out[(size_t)f * nc + j] = val[p];
                if (mask) atomicOr(&mask[(size_t)(j >> 6) * d + f], 1ull << (j & 63));
            }
            else out[(size_t)j * d + f] = val[p];
        }
    }
    if (threadIdx.x == 0) cn[j] = nrm ? nrm[src] : 0.0;
}

// The same transposed block and masks in one launch, one work-group per
// 64-column chunk: lane j of wave 0 reads column j's row range, a wave
// prefix sum places the chunk's nonzeros, and all 256 threads scatter them
// (the column of nonzero e by binary search over the offsets) while the
// mask is built in LDS and then written whole.  Only entries whose mask bit
// is set are written; the CSR row kernels read no others.  d <= kMaskLds.
constexpr int kMaskLds = 4096;
__global__ __launch_bounds__(256) void cols_T_csr_chunk(const int64_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ val,
                                                        const int32_t *__restrict__ idx,
                                                        const double *__restrict__ nrm, int nc, int d,
                                                        double *__restrict__ out, double *__restrict__ cn,
                                                        unsigned long long *__restrict__ mask) {
    __shared__ unsigned long long m[kMaskLds];
    __shared__ int64_t beg[64];
    __shared__ int off[65];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 64;
    const int w = (nc - j0 < 64) ? nc - j0 : 64;
    for (int f = t; f < d; f += 256) m[f] = 0ull;
    if (t < 64) {
        int len = 0;
        int64_t b = 0;
        if (t < w) {
            const int src = idx ? idx[j0 + t] : j0 + t;
            b = rowptr[src];
            len = (int)(rowptr[src + 1] - b);
            cn[j0 + t] = nrm ? nrm[src] : 0.0;
        }
        beg[t] = b;
        int x = len;   // inclusive prefix sum over the wave
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (t >= o) x += y;
        }
        off[t + 1] = x;
        if (t == 0) off[0] = 0;
    }
    __syncthreads();
    const int total = off[64];
    for (int e = t; e < total; e += 256) {
        int lo = 0, hi = 63;   // the last column whose range starts at or before e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const int64_t p = beg[lo] + (e - off[lo]);
        const int f = col[p];
        if ((unsigned)f < (unsigned)d) {
            out[(size_t)f * nc + j0 + lo] = val[p];
            atomicOr(&m[f], 1ull << lo);
        }
    }
    __syncthreads();
    for (int f = t; f < d; f += 256) mask[(size_t)blockIdx.x * d + f] = m[f];
}

// Appends columns j0 .. j0 + gridDim.x - 1 (CSR rows idx[j]) to a
// transposed block of leading dimension ldc: entries, mask bits (ORed into
// the chunk's existing masks) and norms.  One 64-thread work-group per column.
__global__ void cols_append_csr(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                const double *__restrict__ val, const int32_t *__restrict__ idx,
                                const double *__restrict__ nrm, int j0, int d, int ldc, double *__restrict__ XcT,
                                double *__restrict__ cn, unsigned long long *__restrict__ mask) {
    const int j = j0 + blockIdx.x;
    const int src = idx ? idx[j] : j;
    const int64_t p0 = rowptr[src], p1 = rowptr[src + 1];
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const int f = col[p];
        if ((unsigned)f < (unsigned)d) {
            XcT[(size_t)f * ldc + j] = val[p];
            atomicOr(&mask[(size_t)(j >> 6) * d + f], 1ull << (j & 63));
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

// ---- linear kernel: K w = X u with u = sum_j w_j z_j -------------------------
// For k(x, z) = x.z the kernel block never needs forming: two passes over the
// rows' and the columns' features replace rows x cols x d flop by
// (rows + cols) x d.  Fixed summation orders throughout.
constexpr int PRIMAL_SPLITS = 64;   // column ranges summed separately, then in range order

// part[s*d + k] = sum over the columns j of range s of w_j z_j[k] (thread per feature).
__global__ void primal_part_dense(DenseSide cs, int d, const double *__restrict__ w, int per,
                                  double *__restrict__ part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= d) return;
    const int j0 = s * per;
    const int j1 = (j0 + per < cs.n) ? j0 + per : cs.n;
    double acc = 0.0;
    for (int j = j0; j < j1; j++) acc = fma(w[j], cs.X[(size_t)(cs.idx ? cs.idx[j] : j) * cs.ldx + k], acc);
    part[(size_t)s * d + k] = acc;
}

__global__ void primal_sum(const double *__restrict__ part, int nsplit, int d, double *__restrict__ u) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    double s = 0.0;
    for (int p = 0; p < nsplit; p++) s += part[(size_t)p * d + k];
    u[k] = s;
}

__device__ __forceinline__ void primal_store(int i, int gi, double s, double a, double b, double cst,
                                             const double *__restrict__ rsc, double *__restrict__ out) {
    const double sc = rsc ? rsc[i] : 1.0;
    out[gi] = (a == 0.0 ? 0.0 : a * out[gi]) + sc * (b * s) + cst;
}

// out_i = a out_i + rs_i (b x_i.u) + cst, one wave per dense row.
__global__ __launch_bounds__(256) void primal_apply_dense(DenseSide rsd, int d, const double *__restrict__ u, double a,
                                                          double b, double cst, const double *__restrict__ rsc,
                                                          double *__restrict__ out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= rsd.n) return;
    const double *x = rsd.X + (size_t)(rsd.idx ? rsd.idx[i] : i) * rsd.ldx;
    double s = 0.0;
    for (int k = lane; k < d; k += 64) s = fma(x[k], u[k], s);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) primal_store(i, i, s, a, b, cst, rsc, out);
}

// The same for CSR rows: their nonzeros against u.
__global__ __launch_bounds__(256) void primal_apply_csr(const int64_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ val,
                                                        const int32_t *__restrict__ ridx, int n,
                                                        const double *__restrict__ u, double a, double b, double cst,
                                                        const double *__restrict__ rsc, double *__restrict__ out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const int row = ridx ? ridx[i] : i;
    double s = 0.0;
    for (int64_t p = rowptr[row] + lane; p < rowptr[row + 1]; p += 64) s = fma(val[p], u[col[p]], s);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) primal_store(i, i, s, a, b, cst, rsc, out);
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
    if (!c->X && d <= kMaskLds) {
        hipLaunchKernelGGL(cols_T_csr_chunk, dim3((unsigned)nchunk), dim3(256), 0, stream(), c->rowptr, c->col,
                           c->val, c->idx, c->nrm, c->n, d, *XcT, *cn, *mask);
        IRWLS_LAUNCH_CHECK();
        return;
    }
    irwls_dev_memset0(*mask, nchunk * d * sizeof(unsigned long long));
    if (c->X) {
        hipLaunchKernelGGL(cols_T_from_dense, dim3((c->n + 255) / 256, d), dim3(256), 0, stream(), c->X,
                           c->ldx, c->idx, c->nrm, c->n, d, *XcT, *cn, *mask);
    } else {
        // only the entries of features a column has are written (and their
        // mask bits set); the CSR row kernels read an entry only when its
        // bit is set, so the rest of the block needs no clearing
        hipLaunchKernelGGL(cols_from_csr, dim3(c->n), dim3(64), 0, stream(), c->rowptr, c->col, c->val,
                           c->idx, c->nrm, c->n, d, 1, *XcT, *cn, *mask);
    }
    IRWLS_LAUNCH_CHECK();
}

}  // namespace

namespace irwls {

// The CSR row kernels read the columns from a transposed dense block of
// cols x d doubles plus per-chunk feature masks.  For high-dimensional data
// (libsvm text sets with d around 10^6) that block is built and used
// kCsrColBudget bytes of columns at a time, 64 columns at least.
constexpr size_t kCsrColBudget = (size_t)1 << 30;

static int csr_col_chunk(int d) {
    const size_t per = (size_t)(d > 0 ? d : 1) * sizeof(double);
    size_t c = kCsrColBudget / per / 64 * 64;
    if (c < 64) c = 64;
    return c > (size_t)INT32_MAX ? INT32_MAX : (int)c;
}

// CSR rows x a transposed column block (XcT with leading dimension ldc,
// norms, feature masks) of ncols columns: the materialised kernel block.
static void csr_block(const irwls_rows *r, const double *XcT, int ldc, const double *cn,
                      const unsigned long long *cm, int ncols, int d, int kernel, double gamma, double *K, int ldk) {
    const double nz = r->nnz_per_row * r->n;
    TimerScope ts(IRWLS_T_CSR, 2.0 * nz * ncols, 12.0 * nz + 8.0 * ncols * (double)d + 8.0 * r->n * (2.0 + ncols));
    if (ncols <= NARROW) {
        if (r->idx) {
            int glog = 0;
            while ((1 << glog) < ncols) glog++;
            const long threads = (long)r->n << glog;
            ts.launch(kmat_csr_narrow_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, r->rowptr,
                      r->col, r->val, r->nrm, r->idx, r->n, XcT, ldc, d, cm, cn, ncols, glog, kernel, gamma, K, ldk);
        } else {
            ts.launch(kmat_csr_block_narrow_kernel, dim3((r->n + 255) / 256), dim3(256), 0, r->rowptr, r->col,
                      r->val, r->nrm, r->n, XcT, ldc, d, cm, cn, ncols, kernel, gamma, K, ldk);
        }
        return;
    }
    // one wave per (row, 64-column chunk) while the rows alone give few waves
    const int chunks = (ncols + 63) / 64;
    const long waves = (long)r->n * chunks;
    const int gy = r->n >= 8192 ? 1 : (waves <= 65536 ? chunks : (int)((65536 + r->n - 1) / r->n));
    ts.launch(kmat_csr_kernel, dim3((r->n + 3) / 4, gy > 0 ? gy : 1), dim3(256), 0, r->rowptr, r->col, r->val, r->nrm,
              r->idx, r->n, XcT, ldc, d, cm, cn, ncols, kernel, gamma, K, ldk);
}

// Columns [j0, j0 + cn) of a column side, as a side of its own.
static irwls_rows cols_slice(const irwls_rows *c, int j0, int cn) {
    irwls_rows s = *c;
    s.n = cn;
    if (c->idx) {
        s.idx = c->idx + j0;
    } else if (c->X) {
        s.X = c->X + (size_t)j0 * c->ldx;
        s.nrm = c->nrm + j0;
    } else {
        s.rowptr = c->rowptr + j0;
        s.nrm = c->nrm + j0;
    }
    return s;
}

void kmv_empty_cols(int n, double a, double cst, double *out) {
    hipLaunchKernelGGL(axpb_kernel, dim3((n + 255) / 256), dim3(256), 0, stream(), n, a, cst, out);
    IRWLS_LAUNCH_CHECK();
}

// Column splits of the dense kernel-block x vector product for rows x cols.
static void kmv_dense_shape(int rn, int cn, int *rt, int *tps, int *nsplit) {
    *rt = (rn + TB - 1) / TB;
    column_splits(*rt, (cn + TB - 1) / TB, 2, tps, nsplit);
}

size_t kmv_dense_part_bytes(int rn, int cn) {
    int rt, tps, nsplit;
    kmv_dense_shape(rn, cn, &rt, &tps, &nsplit);
    return nsplit > 1 ? (size_t)nsplit * rn * sizeof(double) : 0;
}

size_t kmv_primal_part_bytes(int cn, int d) {
    const int ns = cn < PRIMAL_SPLITS ? cn : PRIMAL_SPLITS;
    return (size_t)(ns + 1) * d * sizeof(double);
}

// Dense rows x dense columns (r->n, c->n > 0, c->X set).  part holds
// kmv_dense_part_bytes(r->n, c->n) bytes (RBF) or kmv_primal_part_bytes(c->n,
// d) (linear kernel, rows dense or CSR).  timed = false records no timer
// events (for callers that may be capturing the stream).
void kmv_dense_cols_run(const irwls_rows *r, const irwls_rows *c, int d, int kernel, double gamma, const double *w,
                        double a, double b, double cst, const double *rs, double *out, double *part, bool timed) {
    const DenseSide csd = dense_side(c);
    if (kernel == 0 && d > 0) {
        const int ns = csd.n < PRIMAL_SPLITS ? csd.n : PRIMAL_SPLITS;
        const int per = (csd.n + ns - 1) / ns;
        double *u = part + (size_t)ns * d;
        const double nz = r->X ? (double)r->n * d : r->nnz_per_row * r->n;
        TimerScope ts(timed ? IRWLS_T_KMV : -1, 2.0 * ((double)csd.n * d + nz),
                      8.0 * ((double)csd.n * d + nz + 2.0 * r->n));
        ts.launch(primal_part_dense, dim3((d + 255) / 256, ns), dim3(256), 0, csd, d, w, per, part);
        ts.launch(primal_sum, dim3((d + 255) / 256), dim3(256), 0, part, ns, d, u);
        if (r->X)
            ts.launch(primal_apply_dense, dim3((r->n + 3) / 4), dim3(256), 0, dense_side(r), d, u, a, b, cst, rs,
                      out);
        else
            ts.launch(primal_apply_csr, dim3((r->n + 3) / 4), dim3(256), 0, r->rowptr, r->col, r->val, r->idx, r->n,
                      u, a, b, cst, rs, out);
        return;
    }
    const DenseSide rsd = dense_side(r);
    int rt, tps, nsplit;
    kmv_dense_shape(rsd.n, csd.n, &rt, &tps, &nsplit);
    const double flops = 2.0 * rsd.n * (double)csd.n * d;
    const double bytes = 8.0 * ((double)rsd.n * d + (double)csd.n * d + 2.0 * rsd.n);
    TimerScope ts(timed ? IRWLS_T_KMV : -1, flops, bytes);
    ts.launch(kmv_dense_kernel, dim3(rt, nsplit), dim3(256), 0, rsd, csd, d, kernel, gamma, w, tps, nsplit, part, a,
              b, cst, rs, out);
    if (nsplit > 1)
        ts.launch(kmv_finish_kernel, dim3((rsd.n + 255) / 256), dim3(256), 0, part, rsd.n, nsplit, a, b, cst, rs, out);
}

}  // namespace irwls

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
        kmv_empty_cols(r->n, a, cst, out);
        return;
    }
    if ((kernel == 0 && d > 0) || r->X) {
        // dense columns (CSR columns scattered into a dense block first); the
        // linear kernel forms u = sum_j w_j z_j, then one dot product per row
        irwls_rows cd = *c;
        if (!c->X) {
            const DenseSide csd = dense_cols(c, d, 6, 7);
            cd.X = csd.X;
            cd.ldx = csd.ldx;
            cd.idx = nullptr;
            cd.nrm = csd.nrm;
        }
        const size_t pb = kmv_part_bytes(r, &cd, d, kernel);
        double *part = pb ? (double *)scratch(kernel == 0 && d > 0 ? 3 : 8, pb) : nullptr;
        kmv_dense_cols_run(r, &cd, d, kernel, gamma, w, a, b, cst, rs, out, part, true);
    } else {
        // column chunks of at most csr_col_chunk(d); the first applies a and
        // cst, the later ones add their sums
        const int chunk = csr_col_chunk(d);
        for (int j0 = 0; j0 < c->n; j0 += chunk) {
            const int cnk = c->n - j0 < chunk ? c->n - j0 : chunk;
            const irwls_rows cs = cols_slice(c, j0, cnk);
            double *XcT, *cn;
            unsigned long long *cm;
            cols_T(&cs, d, &XcT, &cn, &cm);
            // algorithmic: the rows' nonzeros (value + column), their row pointers, the column table, the output
            const double nz = r->nnz_per_row * r->n;
            TimerScope ts(IRWLS_T_CSR, 2.0 * nz * cnk, 12.0 * nz + 8.0 * cnk * (double)d + 24.0 * r->n);
            ts.launch(kmv_csr_kernel, dim3((r->n + 3) / 4), dim3(256), 0, r->rowptr, r->col, r->val, r->nrm, r->idx,
                      r->n, XcT, cnk, d, cm, cn, cnk, kernel, gamma, w + j0, j0 == 0 ? a : 1.0, b,
                      j0 == 0 ? cst : 0.0, rs, out);
        }
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
        ts.launch(kmat_dense_kernel, dim3((r->n + TB - 1) / TB, (c->n + TB - 1) / TB), dim3(256), 0, rsd, csd, d,
                  kernel, gamma, K, ldk);
    } else if (c->n > csr_col_chunk(d)) {
        // high-dimensional CSR: the block in column chunks
        const int chunk = csr_col_chunk(d);
        for (int j0 = 0; j0 < c->n; j0 += chunk) {
            const int cnk = c->n - j0 < chunk ? c->n - j0 : chunk;
            const irwls_rows cs = cols_slice(c, j0, cnk);
            irwls_dev_kernel_matrix(r, &cs, d, kernel, gamma, K + j0, ldk);
        }
    } else {
        double *XcT, *cn;
        unsigned long long *cm;
        cols_T(c, d, &XcT, &cn, &cm);
        csr_block(r, XcT, c->n, cn, cm, c->n, d, kernel, gamma, K, ldk);
    }
}

// ---- growing CSR column sides ----------------------------------------------
struct irwls_coltab {
    double *XcT;                // d x cap, entry (f, j) at f * cap + j (set where the mask bit is)
    double *cn;                 // cap norms
    unsigned long long *mask;   // (cap / 64 chunks) x d feature masks
    int d, cap, n;
};

irwls_coltab *irwls_dev_coltab_create(int d, int cap) {
    if (d <= 0 || cap <= 0 || (size_t)cap * d * sizeof(double) > kCsrColBudget) return nullptr;
    irwls_coltab *t = (irwls_coltab *)calloc(1, sizeof(irwls_coltab));
    if (!t) return nullptr;
    t->d = d;
    t->cap = cap;
    t->n = 0;
    const size_t nchunk = (size_t)(cap + 63) / 64;
    t->XcT = (double *)irwls_dev_alloc((size_t)cap * d * sizeof(double));
    t->cn = (double *)irwls_dev_alloc((size_t)cap * sizeof(double));
    t->mask = (unsigned long long *)irwls_dev_alloc(nchunk * d * sizeof(unsigned long long));   // zero-filled
    return t;
}

void irwls_dev_coltab_destroy(irwls_coltab *t) {
    if (!t) return;
    irwls_dev_free(t->XcT);
    irwls_dev_free(t->cn);
    irwls_dev_free(t->mask);
    free(t);
}

int irwls_dev_coltab_size(const irwls_coltab *t) { return t ? t->n : 0; }

void irwls_dev_coltab_extend(irwls_coltab *t, const irwls_rows *c) {
    if (c->X || c->n > t->cap) {
        fprintf(stderr, "libirwls: irwls_dev_coltab_extend needs CSR rows, at most %d (got %d)\n", t->cap, c->n);
        exit(2);
    }
    if (c->n <= t->n) return;
    hipLaunchKernelGGL(cols_append_csr, dim3(c->n - t->n), dim3(64), 0, stream(), c->rowptr, c->col, c->val, c->idx,
                       c->nrm, t->n, t->d, t->cap, t->XcT, t->cn, t->mask);
    IRWLS_LAUNCH_CHECK();
    t->n = c->n;
}

void irwls_dev_kernel_matrix_coltab(const irwls_rows *r, const irwls_coltab *t, int kernel, double gamma, double *K,
                                    int ldk) {
    if (r->n <= 0 || t->n <= 0) return;
    if (r->X) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matrix_coltab needs CSR rows\n");
        exit(2);
    }
    csr_block(r, t->XcT, t->cap, t->cn, t->mask, t->n, t->d, kernel, gamma, K, ldk);
}

}  // extern "C"
