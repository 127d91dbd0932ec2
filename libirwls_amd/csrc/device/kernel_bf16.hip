// Mixed-precision kernel tiles for gfx950: the dot products of the dense RBF
// and linear kernels (src/kernels.c:63-118) from BF16 hi/lo split operands
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Each feature value x is split into hi = bf16(x) and lo = bf16(x - hi), so
// x = hi + lo to about 16 significant bits, and
//     x.z ~ hi_x.hi_z + hi_x.lo_z + lo_x.hi_z
// (three MFMAs into one FP32 accumulator; the lo.lo term is below that
// precision).  The squared norms are taken from the same representation,
// sum(hi^2 + 2 hi lo), in FP64 and stored in FP32, so |x|^2 + |x|^2 - 2 x.x
// cancels to the accumulation error.  Distances are clamped at zero and the
// exponential is evaluated in FP32; the matvec epilogue sums w_j k_ij in FP64.
//
// Materialised blocks: 128 x 128 tiles per work-group, four waves of 64 x 64
// (4 x 4 MFMA blocks of 16 x 16), K in steps of 32 staged through LDS with
// the next step's loads in flight during the current step's MFMAs.  The
// matvec uses 256 x 256 tiles of eight waves (described at its kernel),
// which halves the operand traffic per product.  MFMA operand
// layout (K = 32): A lane l = row l%16, k = 8(l/16)..8(l/16)+7; B likewise
// with columns; accumulator register r of lane l = C[4(l/16) + r][l%16].
#include "common.h"

using namespace irwls;

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int BT = 128;       // tile rows / columns per work-group
constexpr int KS = 32;        // k step
constexpr int LROW = 40;      // LDS row stride in bf16 (80 bytes: conflict-free fragment reads)

struct Bf16Side {
    const uint16_t *Xh, *Xl;  // rows x ld, row-major
    int ld;                   // multiple of 32
    const float *nrm;         // by original row
    const int32_t *idx;       // may be null
    int n;
};

// Kernel value from a dot product and the two squared norms: the dot itself
// (linear) or 2^(ng2 max(|x|^2 + |z|^2 - 2 x.z, 0)) with ng2 = -gamma log2(e)
// (RBF, one v_exp_f32).  The materialised blocks and the matvec both use this
// function on bit-identical dot products, so the working-set block the
// sub-problem solves with and the error updates see the same kernel values:
// the training loop relies on that (a matvec whose exponential was rounded
// differently from the blocks', by about 1e-7 relatively, took 2.4 times the
// iterations on a batched workload with C up to 100).
__device__ __forceinline__ float kvalue(bool lin, float ng2, float dot, float nr, float nc) {
    float d = fmaf(-2.0f, dot, nr + nc);
    d = d > 0.0f ? d : 0.0f;
    return lin ? dot : __builtin_amdgcn_exp2f(ng2 * d);
}

// Prefetch registers of one k step: thread t holds row t/2, k half (t%2)*16
// of both the row side (A) and the column side (B), hi and lo.
struct StepRegs {
    uint4 ah[2], al[2], bh[2], bl[2];
};

__device__ __forceinline__ void load_side(uint4 (&h)[2], uint4 (&l)[2], const Bf16Side &s, int r0, int k0) {
    const int row = threadIdx.x >> 1, half = threadIdx.x & 1;
    const int g = r0 + row;
    if (g < s.n) {
        const size_t src = (size_t)(s.idx ? s.idx[g] : g) * s.ld + k0 + half * 16;
        const uint4 *ph = reinterpret_cast<const uint4 *>(s.Xh + src);
        const uint4 *pl = reinterpret_cast<const uint4 *>(s.Xl + src);
        h[0] = ph[0];
        h[1] = ph[1];
        l[0] = pl[0];
        l[1] = pl[1];
    } else {
        h[0] = h[1] = l[0] = l[1] = make_uint4(0, 0, 0, 0);
    }
}

__device__ __forceinline__ void store_side(uint16_t (*S)[BT][LROW], const uint4 (&h)[2], const uint4 (&l)[2]) {
    const int row = threadIdx.x >> 1, half = threadIdx.x & 1;
    uint4 *dh = reinterpret_cast<uint4 *>(&S[0][row][half * 16]);
    uint4 *dl = reinterpret_cast<uint4 *>(&S[1][row][half * 16]);
    dh[0] = h[0];
    dh[1] = h[1];
    dl[0] = l[0];
    dl[1] = l[1];
}

__device__ __forceinline__ v8bf frag(const uint16_t (*S)[LROW], int row0) {
    const int lane = threadIdx.x & 63;
    const uint4 v = *reinterpret_cast<const uint4 *>(&S[row0 + (lane & 15)][8 * (lane >> 4)]);
    return __builtin_bit_cast(v8bf, v);
}

// acc[mb][nb] = dot products of the wave's 64 x 64 block of the tile
// (rows r0 + 64 wr + 16 mb + 4(l/16) + r, cols c0 + 64 wc + 16 nb + l%16).
__device__ __forceinline__ void tile_dots_bf16(v4f (&acc)[4][4], uint16_t (*As)[BT][LROW], uint16_t (*Bs)[BT][LROW],
                                               const Bf16Side &rs, const Bf16Side &cs, int r0, int c0, int kp) {
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[mb][nb] = (v4f){0.f, 0.f, 0.f, 0.f};
    StepRegs pf;
    load_side(pf.ah, pf.al, rs, r0, 0);
    load_side(pf.bh, pf.bl, cs, c0, 0);
    for (int k0 = 0; k0 < kp; k0 += KS) {
        __syncthreads();   // the previous step's fragments are read
        store_side(As, pf.ah, pf.al);
        store_side(Bs, pf.bh, pf.bl);
        __syncthreads();
        if (k0 + KS < kp) {
            load_side(pf.ah, pf.al, rs, r0, k0 + KS);
            load_side(pf.bh, pf.bl, cs, c0, k0 + KS);
        }
        v8bf ah[4], al[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            ah[mb] = frag(As[0], wr * 64 + mb * 16);
            al[mb] = frag(As[1], wr * 64 + mb * 16);
        }
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            const v8bf bh = frag(Bs[0], wc * 64 + nb * 16);
            const v8bf bl = frag(Bs[1], wc * 64 + nb * 16);
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mb], bh, acc[mb][nb], 0, 0, 0);
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mb], bl, acc[mb][nb], 0, 0, 0);
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mb], bh, acc[mb][nb], 0, 0, 0);
            }
        }
    }
}

// ---- kernel block x vector: 256 x 256 tiles --------------------------------
// Eight waves per work-group, each a 128 (columns, MFMA A side) x 64 (rows,
// B side) block, so a lane keeps four fixed rows and its per-row sums stay in
// registers.  Both sides of a k step (hi and lo planes, 64 bytes per row)
// are staged in double-buffered LDS without padding: the four 16-byte
// chunks of row r are stored in the order q ^ (((r >> 1) ^ (r >> 3)) & 3),
// so the 16 rows of a fragment read fall on distinct banks.  The next k
// step's global loads (across tile boundaries) are in flight during the
// current step's MFMAs.
constexpr int MT = 256;

__device__ __forceinline__ int chunk_swz(int row, int q) { return q ^ (((row >> 1) ^ (row >> 3)) & 3); }

// Thread t's share of one k step of a 256-row side: row t/2, k half t%2 of
// both planes (zeros past the last row).
__device__ __forceinline__ void load_step(uint4 (&h)[2], uint4 (&l)[2], const Bf16Side &s, int r0, int k0) {
    const int row = threadIdx.x >> 1, half = threadIdx.x & 1;
    const int g = r0 + row;
    if (g < s.n) {
        const size_t src = (size_t)(s.idx ? s.idx[g] : g) * s.ld + k0 + half * 16;
        const uint4 *ph = reinterpret_cast<const uint4 *>(s.Xh + src);
        const uint4 *pl = reinterpret_cast<const uint4 *>(s.Xl + src);
        h[0] = ph[0];
        h[1] = ph[1];
        l[0] = pl[0];
        l[1] = pl[1];
    } else {
        h[0] = h[1] = l[0] = l[1] = make_uint4(0, 0, 0, 0);
    }
}

// S: hi plane [MT][32] followed by the lo plane.
__device__ __forceinline__ void store_step(uint16_t *S, const uint4 (&h)[2], const uint4 (&l)[2]) {
    const int row = threadIdx.x >> 1, half = threadIdx.x & 1;
    uint16_t *rh = S + row * 32, *rl = S + (MT + row) * 32;
    *reinterpret_cast<uint4 *>(rh + 8 * chunk_swz(row, 2 * half)) = h[0];
    *reinterpret_cast<uint4 *>(rh + 8 * chunk_swz(row, 2 * half + 1)) = h[1];
    *reinterpret_cast<uint4 *>(rl + 8 * chunk_swz(row, 2 * half)) = l[0];
    *reinterpret_cast<uint4 *>(rl + 8 * chunk_swz(row, 2 * half + 1)) = l[1];
}

__device__ __forceinline__ v8bf frag_swz(const uint16_t *P, int row0) {
    const int lane = threadIdx.x & 63;
    const int row = row0 + (lane & 15);
    const uint4 v = *reinterpret_cast<const uint4 *>(P + row * 32 + 8 * chunk_swz(row, lane >> 4));
    return __builtin_bit_cast(v8bf, v);
}

__device__ __forceinline__ float side_nrm(const Bf16Side &s, int i) {
    return s.nrm[s.idx ? s.idx[i] : i];
}

// grid (row tiles, column splits), split s covers column tiles
// [s*tps, (s+1)*tps).  Kernel values as kvalue(); each lane adds w_j k_ij to
// its row sums in FP64 (FP64 w times the FP32 kernel value), so the training
// loop's error updates e -= K dbeta add up exactly to e0 - K beta for the
// fixed kernel values.  The three products of a k step are issued in the
// materialised blocks' order (hi.hi, lo of the rows . hi of the columns,
// hi of the rows . lo of the columns) with the operands' roles swapped,
// which gives the same FP32 dot products.  Partial sums to part[s*n + i]
// when nsplit > 1, else the output epilogue out = a out + rs (b sum) + cst.
__global__ __launch_bounds__(512) void kmv_bf16_kernel(Bf16Side rs, Bf16Side cs, int kp, int kernel, float ng2,
                                                       const double *__restrict__ w, int tps, int nsplit,
                                                       double *__restrict__ part, double a, double b, double cst,
                                                       const double *__restrict__ rsc, double *__restrict__ out) {
    __shared__ __align__(16) uint16_t L[2][2][2 * MT * 32];   // [buffer][A columns, B rows][hi plane, lo plane]
    __shared__ __align__(16) double wt[3][MT];                // [tile % 3][column]: w
    __shared__ __align__(16) float nt[3][MT];                 // [tile % 3][column]: column norm
    __shared__ double red[2][MT];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wa = wave >> 2, wb = wave & 3;
    const int r0 = blockIdx.x * MT;
    const int ntiles = (cs.n + MT - 1) / MT;
    const int ct_begin = blockIdx.y * tps;
    const int ct_end = (ct_begin + tps < ntiles) ? ct_begin + tps : ntiles;
    const int nks = kp / 32;
    const int total = (ct_end - ct_begin) * nks;   // (tile, k step) pairs of this work-group
    const bool lin = kernel == 0;                  // linear kernel: k = x.z (a select, no branch)
    float nr[4];
    double rsum[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int gi = r0 + wb * 64 + nb * 16 + (lane & 15);
        nr[nb] = (gi < rs.n) ? side_nrm(rs, gi) : 0.0f;
        rsum[nb] = 0.0;
    }
    // w and the column norms of tile t; three buffers, so the one written is
    // never the one a slower wave may still read in the previous epilogue
    auto stage_cols = [&](int t) {
        if (threadIdx.x < MT) {
            const int gj = (ct_begin + t) * MT + threadIdx.x;
            const bool in = gj < cs.n;
            wt[t % 3][threadIdx.x] = in ? w[gj] : 0.0;
            nt[t % 3][threadIdx.x] = in ? side_nrm(cs, gj) : 0.0f;
        }
    };
    if (total > 0) {
        uint4 ah[2], al[2], bh[2], bl[2];
        load_step(ah, al, cs, ct_begin * MT, 0);
        load_step(bh, bl, rs, r0, 0);
        stage_cols(0);
        store_step(L[0][0], ah, al);
        store_step(L[0][1], bh, bl);
        __syncthreads();
        v4f acc[8][4];
#pragma unroll
        for (int mb = 0; mb < 8; mb++)
#pragma unroll
            for (int nb = 0; nb < 4; nb++) acc[mb][nb] = (v4f){0.f, 0.f, 0.f, 0.f};
        int t = 0, s = 0;
        for (int it = 0; it < total; it++) {
            const int cur = it & 1;
            const bool more = it + 1 < total;
            int tn = t, sn = s + 1;
            if (sn == nks) {
                tn = t + 1;
                sn = 0;
            }
            if (more) {
                load_step(ah, al, cs, (ct_begin + tn) * MT, sn * 32);
                load_step(bh, bl, rs, r0, sn * 32);
            }
            const uint16_t *PA = L[cur][0], *PB = L[cur][1];
            v8bf fbh[4], fbl[4];
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                fbh[nb] = frag_swz(PB, wb * 64 + nb * 16);
                fbl[nb] = frag_swz(PB + MT * 32, wb * 64 + nb * 16);
            }
#pragma unroll
            for (int mb = 0; mb < 8; mb++) {
                const v8bf fah = frag_swz(PA, wa * 128 + mb * 16);
                const v8bf fal = frag_swz(PA + MT * 32, wa * 128 + mb * 16);
#pragma unroll
                for (int nb = 0; nb < 4; nb++) {
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah, fbh[nb], acc[mb][nb], 0, 0, 0);
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fal, fbh[nb], acc[mb][nb], 0, 0, 0);
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah, fbl[nb], acc[mb][nb], 0, 0, 0);
                }
            }
            if (more) {
                // the other buffer was last read in the previous step, before its barrier
                store_step(L[cur ^ 1][0], ah, al);
                store_step(L[cur ^ 1][1], bh, bl);
                if (sn == 0) stage_cols(tn);
            }
            __syncthreads();
            if (s == nks - 1) {
                // epilogue of tile t: acc[mb][nb][r] is the dot product of column
                // wa*128 + mb*16 + 4 (l/16) + r and row wb*64 + nb*16 + l%16
                const int wbuf = t % 3;
#pragma unroll
                for (int mb = 0; mb < 8; mb++) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int cl = wa * 128 + mb * 16 + 4 * (lane >> 4) + r;
                        const double wj = wt[wbuf][cl];
                        const float nc = nt[wbuf][cl];
#pragma unroll
                        for (int nb = 0; nb < 4; nb++)
                            rsum[nb] = fma(wj, (double)kvalue(lin, ng2, acc[mb][nb][r], nr[nb], nc), rsum[nb]);
                    }
                    __builtin_amdgcn_sched_barrier(0);   // keep each block's loads and temporaries local
                }
#pragma unroll
                for (int mb = 0; mb < 8; mb++)
#pragma unroll
                    for (int nb = 0; nb < 4; nb++) acc[mb][nb] = (v4f){0.f, 0.f, 0.f, 0.f};
            }
            t = tn;
            s = sn;
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
    if (threadIdx.x < MT) {
        const int i = threadIdx.x;
        const int gi = r0 + i;
        if (gi < rs.n) {
            const double sv = red[0][i] + red[1][i];
            if (nsplit > 1) {
                part[(size_t)blockIdx.y * rs.n + gi] = sv;
            } else {
                const double sc = rsc ? rsc[gi] : 1.0;
                out[gi] = (a == 0.0 ? 0.0 : a * out[gi]) + sc * (b * sv) + cst;
            }
        }
    }
}

// Materialised kernel block K[i*ldk + j] (FP64 storage of the FP32 values).
__global__ __launch_bounds__(256) void kmat_bf16_kernel(Bf16Side rs, Bf16Side cs, int kp, int kernel, float ng2,
                                                        double *__restrict__ K, int ldk) {
    __shared__ __align__(16) uint16_t As[2][BT][LROW];
    __shared__ __align__(16) uint16_t Bs[2][BT][LROW];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = blockIdx.x * BT, c0 = blockIdx.y * BT;
    v4f acc[4][4];
    tile_dots_bf16(acc, As, Bs, rs, cs, r0, c0, kp);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int gj = c0 + wc * 64 + nb * 16 + (lane & 15);
        if (gj >= cs.n) continue;
        const float ncj = side_nrm(cs, gj);
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gi = r0 + wr * 64 + mb * 16 + 4 * (lane >> 4) + r;
                if (gi < rs.n)
                    K[(size_t)gi * ldk + gj] = kvalue(kernel == 0, ng2, acc[mb][nb][r], side_nrm(rs, gi), ncj);
            }
    }
}

__global__ void kmv_finish_bf16(const double *__restrict__ part, int n, int nsplit, double a, double b, double cst,
                                const double *__restrict__ rsc, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; k++) s += part[(size_t)k * n + i];
    const double sc = rsc ? rsc[i] : 1.0;
    out[i] = (a == 0.0 ? 0.0 : a * out[i]) + sc * (b * s) + cst;
}

// hi/lo split of row-major FP64 rows; columns d..ldb-1 are zero.  One thread
// per row writes its norm sum(hi^2 + 2 hi lo) (FP64, stored FP32).
__global__ void split_bf16_kernel(const double *__restrict__ X, int ldx, int n, int d, uint16_t *__restrict__ Xh,
                                  uint16_t *__restrict__ Xl, int ldb) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(e / ldb), k = (int)(e % ldb);
    if (i >= n) return;
    uint16_t hb = 0, lb = 0;
    if (k < d) {
        const double x = X[(size_t)i * ldx + k];
        const __bf16 h = (__bf16)(float)x;
        const double lo = x - (double)(float)h;
        const __bf16 l = (__bf16)(float)lo;
        hb = __builtin_bit_cast(uint16_t, h);
        lb = __builtin_bit_cast(uint16_t, l);
    }
    Xh[(size_t)i * ldb + k] = hb;
    Xl[(size_t)i * ldb + k] = lb;
}

// One wave per row: lane l sums features l, l + 64, ... in FP64, then the
// wave's lanes are combined in a fixed butterfly order.
__global__ __launch_bounds__(256) void split_norms_kernel(const uint16_t *__restrict__ Xh,
                                                          const uint16_t *__restrict__ Xl, int ldb, int n,
                                                          float *__restrict__ nrm) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    double s = 0.0;
    for (int k = lane; k < ldb; k += 64) {
        const double h = (double)(float)__builtin_bit_cast(__bf16, Xh[(size_t)i * ldb + k]);
        const double l = (double)(float)__builtin_bit_cast(__bf16, Xl[(size_t)i * ldb + k]);
        s += h * h + 2.0 * h * l;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) nrm[i] = (float)s;
}

// ng2 of kvalue() for a kernel width gamma (one rounding, shared by every launch).
float exp2_scale(double gamma) {
    return (float)(-gamma * 1.4426950408889634);
}

Bf16Side side_of(const irwls_rows_bf16 *r) {
    Bf16Side s;
    s.Xh = r->Xh;
    s.Xl = r->Xl;
    s.ld = r->ld;
    s.nrm = r->nrm;
    s.idx = r->idx;
    s.n = r->n;
    return s;
}

}  // namespace

extern "C" {

void irwls_dev_bf16_split(const double *X, int ldx, int n, int d, uint16_t *Xh, uint16_t *Xl, int ldb, float *nrm) {
    if (n <= 0) return;
    if (ldb % KS != 0 || ldb < d) {
        fprintf(stderr, "libirwls: irwls_dev_bf16_split needs ldb >= d and a multiple of %d (d=%d, ldb=%d)\n", KS, d,
                ldb);
        exit(2);
    }
    const size_t total = (size_t)n * ldb;
    hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream(), X, ldx, n, d,
                       Xh, Xl, ldb);
    IRWLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(split_norms_kernel, dim3((n + 3) / 4), dim3(256), 0, stream(), Xh, Xl, ldb, n, nrm);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_kernel_matvec_bf16(const irwls_rows_bf16 *r, const irwls_rows_bf16 *c, int kernel, double gamma,
                                  const double *w, double a, double b, double cst, const double *rs, double *out) {
    if (r->n <= 0) return;
    if (r->ld != c->ld) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matvec_bf16: row sides of different widths\n");
        exit(2);
    }
    if (c->n <= 0) {
        // no columns: out = a out + cst
        irwls_rows rr = {}, none = {};
        rr.n = r->n;
        irwls_dev_kernel_matvec(&rr, &none, 0, kernel, gamma, w, a, b, cst, rs, out);
        return;
    }
    const int rt = (r->n + MT - 1) / MT;
    const int ct = (c->n + MT - 1) / MT;
    int tps, nsplit;
    column_splits(rt, ct, 1, &tps, &nsplit);
    double *part = nullptr;
    if (nsplit > 1) part = (double *)scratch(14, (size_t)nsplit * r->n * sizeof(double));
    const double flops = 2.0 * r->n * (double)c->n * r->ld;
    const double bytes = 4.0 * ((double)r->n * r->ld + (double)c->n * r->ld) + 8.0 * 2.0 * r->n;
    TimerScope ts(IRWLS_T_BF16, flops, bytes);
    ts.launch(kmv_bf16_kernel, dim3(rt, nsplit), dim3(512), 0, side_of(r), side_of(c), r->ld, kernel,
              exp2_scale(gamma), w, tps, nsplit, part, a, b, cst, rs, out);
    if (nsplit > 1)
        ts.launch(kmv_finish_bf16, dim3((r->n + 255) / 256), dim3(256), 0, part, r->n, nsplit, a, b, cst, rs, out);
}

void irwls_dev_kernel_matrix_bf16(const irwls_rows_bf16 *r, const irwls_rows_bf16 *c, int kernel, double gamma,
                                  double *K, int ldk) {
    if (r->n <= 0 || c->n <= 0) return;
    if (r->ld != c->ld) {
        fprintf(stderr, "libirwls: irwls_dev_kernel_matrix_bf16: row sides of different widths\n");
        exit(2);
    }
    const double flops = 2.0 * r->n * (double)c->n * r->ld;
    const double bytes = 4.0 * ((double)r->n * r->ld + (double)c->n * r->ld) + 8.0 * (double)r->n * c->n;
    TimerScope ts(IRWLS_T_BF16, flops, bytes);
    ts.launch(kmat_bf16_kernel, dim3((r->n + BT - 1) / BT, (c->n + BT - 1) / BT), dim3(256), 0, side_of(r),
              side_of(c), r->ld, kernel, exp2_scale(gamma), K, ldk);
}

}  // extern "C"
