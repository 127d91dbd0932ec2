// Proposal network for SGMA candidates: a 32 -> 128 -> 128 -> 1 MLP with
// ReLU hidden layers.  Training in FP32 (fused forward + backward per block
// of 64 samples, deterministic reduction of the per-block gradients, Adam);
// inference in BF16 on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Parameter vector (FP32, MLP_NPARAMS entries): W1[32][128], b1[128],
// W2[128][128], b2[128], w3[128], b3 (W[i][j] = weight from input i to
// unit j).  Loss: mean squared error over the batch.
#include "common.h"

using namespace irwls;

namespace {

constexpr int NIN = 32, NH = 128;
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH;
constexpr int NP = OB3 + 1;
constexpr int SB = 64;        // samples per work-group
constexpr int HLD = NH + 1;   // padded LDS row

static_assert(NP == IRWLS_MLP_NPARAMS, "parameter layout");

// Fused forward and backward pass of up to 64 samples; writes this block's
// gradient sum (of the mean-squared-error loss over nb samples) to
// gpart[blockIdx.x * NP ...] and its squared-error sum to lpart[blockIdx.x].
__global__ __launch_bounds__(256) void mlp_fwd_bwd(const float *__restrict__ P, const float *__restrict__ X,
                                                   const float *__restrict__ T, int nb, float *__restrict__ gpart,
                                                   float *__restrict__ lpart) {
    __shared__ float xs[SB][NIN + 1];
    __shared__ float h1[SB][HLD];
    __shared__ float h2[SB][HLD];
    __shared__ float dh[SB][HLD];
    __shared__ float dy[SB];
    const int t = threadIdx.x;
    const int s0 = blockIdx.x * SB;
    const int nv = (nb - s0 < SB) ? nb - s0 : SB;
    float *g = gpart + (size_t)blockIdx.x * NP;
    for (int e = t; e < SB * NIN; e += 256) {
        const int s = e / NIN, i = e % NIN;
        xs[s][i] = (s < nv) ? X[(size_t)(s0 + s) * NIN + i] : 0.0f;
    }
    __syncthreads();
    const int j = t & (NH - 1), sh = (t >> 7) * (SB / 2);   // unit j, samples sh..sh+31
    // layer 1 (each weight loaded once for the thread's 32 samples; every
    // sample's sum still runs over i in order)
    float acc[SB / 2];
    {
        const float b = P[OB1 + j];
#pragma unroll
        for (int u = 0; u < SB / 2; u++) acc[u] = b;
        for (int i = 0; i < NIN; i++) {
            const float w = P[OW1 + i * NH + j];
#pragma unroll
            for (int u = 0; u < SB / 2; u++) acc[u] = fmaf(xs[sh + u][i], w, acc[u]);
        }
#pragma unroll
        for (int u = 0; u < SB / 2; u++) h1[sh + u][j] = acc[u] > 0.0f ? acc[u] : 0.0f;
    }
    __syncthreads();
    // layer 2
    {
        const float b = P[OB2 + j];
#pragma unroll
        for (int u = 0; u < SB / 2; u++) acc[u] = b;
        for (int i = 0; i < NH; i++) {
            const float w = P[OW2 + i * NH + j];
#pragma unroll
            for (int u = 0; u < SB / 2; u++) acc[u] = fmaf(h1[sh + u][i], w, acc[u]);
        }
#pragma unroll
        for (int u = 0; u < SB / 2; u++) h2[sh + u][j] = acc[u] > 0.0f ? acc[u] : 0.0f;
    }
    __syncthreads();
    // output and loss
    if (t < SB) {
        float y = P[OB3];
        for (int i = 0; i < NH; i++) y = fmaf(h2[t][i], P[OW3 + i], y);
        const float err = (t < nv) ? y - T[s0 + t] : 0.0f;
        dy[t] = 2.0f * err / (float)nb;
        dh[t][0] = err * err;   // scratch for the loss sum
    }
    __syncthreads();
    if (t == 0) {
        float l = 0.0f;
        for (int s = 0; s < SB; s++) l += dh[s][0];
        lpart[blockIdx.x] = l;
    }
    __syncthreads();
    // w3, b3 and dh2
    if (t < NH) {
        float a = 0.0f;
        for (int s = 0; s < SB; s++) a = fmaf(dy[s], h2[s][t], a);
        g[OW3 + t] = a;
    }
    if (t == 0) {
        float a = 0.0f;
        for (int s = 0; s < SB; s++) a += dy[s];
        g[OB3] = a;
    }
    __syncthreads();
    for (int s = sh; s < sh + SB / 2; s++) dh[s][j] = (h2[s][j] > 0.0f) ? dy[s] * P[OW3 + j] : 0.0f;
    __syncthreads();
    // W2, b2
    for (int e = t; e < NH * NH; e += 256) {
        const int i = e / NH, jj = e % NH;
        float a = 0.0f;
        for (int s = 0; s < SB; s++) a = fmaf(h1[s][i], dh[s][jj], a);
        g[OW2 + e] = a;
    }
    if (t < NH) {
        float a = 0.0f;
        for (int s = 0; s < SB; s++) a += dh[s][t];
        g[OB2 + t] = a;
    }
    // dh1 (into h2, no longer needed): sum_j dh2[s][j] W2[i][j], masked by h1 > 0
#pragma unroll
    for (int u = 0; u < SB / 2; u++) acc[u] = 0.0f;
    for (int jj = 0; jj < NH; jj++) {
        const float w = P[OW2 + j * NH + jj];
#pragma unroll
        for (int u = 0; u < SB / 2; u++) acc[u] = fmaf(dh[sh + u][jj], w, acc[u]);
    }
#pragma unroll
    for (int u = 0; u < SB / 2; u++) h2[sh + u][j] = (h1[sh + u][j] > 0.0f) ? acc[u] : 0.0f;
    __syncthreads();
    // W1, b1
    for (int e = t; e < NIN * NH; e += 256) {
        const int i = e / NH, jj = e % NH;
        float a = 0.0f;
        for (int s = 0; s < SB; s++) a = fmaf(xs[s][i], h2[s][jj], a);
        g[OW1 + e] = a;
    }
    if (t < NH) {
        float a = 0.0f;
        for (int s = 0; s < SB; s++) a += h2[s][t];
        g[OB1 + t] = a;
    }
}

// grad[p] = sum over blocks (fixed order); loss[0] = mean squared error.
// Eight blocks' loads are issued before their eight additions, made in
// block order, so each thread keeps eight reads in flight.
__global__ void mlp_reduce(const float *__restrict__ gpart, const float *__restrict__ lpart, int nblk, int nb,
                           float *__restrict__ grad, float *__restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NP) {
        float a = 0.0f;
        int b = 0;
        for (; b + 8 <= nblk; b += 8) {
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = gpart[(size_t)(b + q) * NP + p];
#pragma unroll
            for (int q = 0; q < 8; q++) a += v[q];
        }
        for (; b < nblk; b++) a += gpart[(size_t)b * NP + p];
        grad[p] = a;
    }
    if (p == 0 && loss) {
        float l = 0.0f;
        for (int b = 0; b < nblk; b++) l += lpart[b];
        loss[0] = l / (float)nb;
    }
}

__global__ void mlp_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m,
                         float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    const float g = G[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// FP32 forward of nb samples, one thread per sample (reference path for
// checks; inference uses the BF16 kernel).
__global__ void mlp_forward_f32(const float *__restrict__ P, const float *__restrict__ X, int nb,
                                float *__restrict__ Y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    float h1v[NH];
    for (int j = 0; j < NH; j++) {
        float a = P[OB1 + j];
        for (int i = 0; i < NIN; i++) a = fmaf(X[(size_t)s * NIN + i], P[OW1 + i * NH + j], a);
        h1v[j] = a > 0.0f ? a : 0.0f;
    }
    float y = P[OB3];
    for (int j = 0; j < NH; j++) {
        float a = P[OB2 + j];
        for (int i = 0; i < NH; i++) a = fmaf(h1v[i], P[OW2 + i * NH + j], a);
        y = fmaf(a > 0.0f ? a : 0.0f, P[OW3 + j], y);
    }
    Y[s] = y;
}

// ---------------------------------------------------------------- BF16 ---
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Packed inference weights: W1T[128][32] and W2T[128][128] in BF16
// (transposed so that the 8 consecutive k of an MFMA B fragment are
// contiguous).
constexpr int PK_W1T = 0, PK_W2T = NH * NIN, PK_N = PK_W2T + NH * NH;

__global__ void mlp_pack(const float *__restrict__ P, uint16_t *__restrict__ Q) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < NH * NIN) {
        const int j = e / NIN, i = e % NIN;   // W1T[j][i] = W1[i][j]
        Q[PK_W1T + e] = __builtin_bit_cast(uint16_t, (__bf16)P[OW1 + i * NH + j]);
    } else if (e < PK_N) {
        const int f = e - PK_W2T;
        const int j = f / NH, i = f % NH;     // W2T[j][i] = W2[i][j]
        Q[e] = __builtin_bit_cast(uint16_t, (__bf16)P[OW2 + i * NH + j]);
    }
}

// 64 samples per work-group; wave w computes rows 16w..16w+15 of each layer
// (8 column blocks of 16 units).  A fragment: lane l = row l%16, k
// 8(l/16)..+7; B fragment: lane l = unit l%16, same k; accumulator register r
// of lane l = (row 4(l/16) + r, unit l%16).
__global__ __launch_bounds__(256) void mlp_infer_bf16(const uint16_t *__restrict__ Q, const float *__restrict__ P,
                                                      const float *__restrict__ X, int nb, float *__restrict__ Y) {
    __shared__ __align__(16) uint16_t xb[SB][NIN + 8];
    __shared__ __align__(16) uint16_t hb[SB][NH + 8];
    __shared__ float yp[SB][NH / 16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int s0 = blockIdx.x * SB;
    for (int e = t; e < SB * NIN; e += 256) {
        const int s = e / NIN, i = e % NIN;
        const float v = (s0 + s < nb) ? X[(size_t)(s0 + s) * NIN + i] : 0.0f;
        xb[s][i] = __builtin_bit_cast(uint16_t, (__bf16)v);
    }
    __syncthreads();
    const int row0 = wave * 16;
    // layer 1: K = 32, one MFMA per 16 x 16 block
    v4f acc[8];
    {
        const v8bf a = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&xb[row0 + (lane & 15)][8 * (lane >> 4)]));
#pragma unroll
        for (int cb = 0; cb < 8; cb++) {
            const int unit = cb * 16 + (lane & 15);
            const v8bf b = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&Q[PK_W1T + unit * NIN + 8 * (lane >> 4)]));
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, (v4f){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 8; cb++) {
        const int unit = cb * 16 + (lane & 15);
        const float bias = P[OB1 + unit];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float v = acc[cb][r] + bias;
            hb[row0 + 4 * (lane >> 4) + r][unit] = __builtin_bit_cast(uint16_t, (__bf16)(v > 0.0f ? v : 0.0f));
        }
    }
    __syncthreads();
    // layer 2: K = 128 in four steps
#pragma unroll
    for (int cb = 0; cb < 8; cb++) acc[cb] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NH / 32; ks++) {
        const v8bf a = __builtin_bit_cast(
            v8bf, *reinterpret_cast<const uint4 *>(&hb[row0 + (lane & 15)][ks * 32 + 8 * (lane >> 4)]));
#pragma unroll
        for (int cb = 0; cb < 8; cb++) {
            const int unit = cb * 16 + (lane & 15);
            const v8bf b = __builtin_bit_cast(
                v8bf, *reinterpret_cast<const uint4 *>(&Q[PK_W2T + unit * NH + ks * 32 + 8 * (lane >> 4)]));
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[cb], 0, 0, 0);
        }
    }
    // output: relu(h2) . w3, FP32
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int cb = 0; cb < 8; cb++) {
        const int unit = cb * 16 + (lane & 15);
        const float bias = P[OB2 + unit], w3 = P[OW3 + unit];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float v = acc[cb][r] + bias;
            part[r] = fmaf(v > 0.0f ? v : 0.0f, w3, part[r]);
        }
    }
    // sum over the 16 units of the lane group
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float v = part[r];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        part[r] = v;
    }
    if ((lane & 15) == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) yp[row0 + 4 * (lane >> 4) + r][0] = part[r];
    }
    __syncthreads();
    if (t < SB && s0 + t < nb) Y[s0 + t] = yp[t][0] + P[OB3];
}

// ------------------------------------------------- candidate features ---
// 32 features per candidate c of a pool of P (one work-group; P <= 1024):
// its label, the novelty eta_c = 1 - k_C(c)^T K_C^-1 k_C(c), the squared
// residual of its kernel column on a fixed row subsample (gsub, and gsub2 on
// the subsample's first half), the estimated gain log(gsub/eta) raw and
// relative to the pool, statistics of its kernel values to the current
// centroids (KNC row), |x_c|^2 and the step.
constexpr int FMAXP = 1024;

// Kernel values of each candidate to the current centroids (its KNC row,
// s entries), summarised by one wave per candidate with coalesced reads:
// max, min, sum, sum of squares and the counts above 0.1, 0.5 and 0.9, each
// lane over every 64th entry and the lanes combined in a fixed butterfly.
__global__ __launch_bounds__(256) void knc_stats_kernel(int P, int s, const double *__restrict__ KNC, int ldk,
                                                        float *__restrict__ st) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= P) return;
    float kmax = 0.f, kmin = (s > 0) ? 1e30f : 0.f, ksum = 0.f, ksq = 0.f, c1 = 0.f, c5 = 0.f, c9 = 0.f;
    for (int e = lane; e < s; e += 64) {
        const float k = (float)KNC[(size_t)c * ldk + e];
        kmax = fmaxf(kmax, k);
        kmin = fminf(kmin, k);
        ksum += k;
        ksq += k * k;
        c1 += (k > 0.1f) ? 1.f : 0.f;
        c5 += (k > 0.5f) ? 1.f : 0.f;
        c9 += (k > 0.9f) ? 1.f : 0.f;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        kmax = fmaxf(kmax, __shfl_xor(kmax, off));
        kmin = fminf(kmin, __shfl_xor(kmin, off));
        ksum += __shfl_xor(ksum, off);
        ksq += __shfl_xor(ksq, off);
        c1 += __shfl_xor(c1, off);
        c5 += __shfl_xor(c5, off);
        c9 += __shfl_xor(c9, off);
    }
    if (lane == 0) {
        float *o = st + (size_t)c * 8;
        o[0] = kmax;
        o[1] = kmin;
        o[2] = ksum;
        o[3] = ksq;
        o[4] = c1;
        o[5] = c5;
        o[6] = c9;
    }
}

__global__ __launch_bounds__(256) void sgma_features_kernel(int P, int s, int B, const double *__restrict__ KNC,
                                                            int ldk, const double *__restrict__ eta,
                                                            const double *__restrict__ gsub,
                                                            const double *__restrict__ gsub2,
                                                            const double *__restrict__ ycand,
                                                            const double *__restrict__ qvcand, float pos_frac,
                                                            const int32_t *__restrict__ npos,
                                                            const float *__restrict__ kst, float *__restrict__ F) {
    __shared__ float est[FMAXP], leta[FMAXP], lg[FMAXP];
    __shared__ float red[6][256];
    const int t = threadIdx.x;
    float se = 0.f, se2 = 0.f, smax = -1e30f, smin = 1e30f, sle = 0.f, slg = 0.f;
    for (int c = t; c < P; c += 256) {
        const double e = eta[c] > 1e-12 ? eta[c] : 1e-12;
        const double g = gsub[c] > 1e-30 ? gsub[c] : 1e-30;
        est[c] = (float)(log(g) - log(e));
        leta[c] = (float)log(e);
        lg[c] = (float)log(g);
        se += est[c];
        se2 += est[c] * est[c];
        smax = fmaxf(smax, est[c]);
        smin = fminf(smin, est[c]);
        sle += leta[c];
        slg += lg[c];
    }
    red[0][t] = se;
    red[1][t] = se2;
    red[2][t] = smax;
    red[3][t] = smin;
    red[4][t] = sle;
    red[5][t] = slg;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (t < w) {
            red[0][t] += red[0][t + w];
            red[1][t] += red[1][t + w];
            red[2][t] = fmaxf(red[2][t], red[2][t + w]);
            red[3][t] = fminf(red[3][t], red[3][t + w]);
            red[4][t] += red[4][t + w];
            red[5][t] += red[5][t + w];
        }
        __syncthreads();
    }
    const float inv = 1.0f / (float)(P > 0 ? P : 1);
    const float mean = red[0][0] * inv;
    const float sd = sqrtf(fmaxf(red[1][0] * inv - mean * mean, 1e-12f));
    const float emax = red[2][0], emin = red[3][0];
    const float mleta = red[4][0] * inv, mlg = red[5][0] * inv;
    for (int c = t; c < P; c += 256) {
        const float *ks = kst + (size_t)c * 8;
        const float kmax = ks[0], kmin = ks[1], ksum = ks[2], ksq = ks[3], c1 = ks[4], c5 = ks[5], c9 = ks[6];
        const float is = 1.0f / (float)(s > 0 ? s : 1);
        const float y = (float)ycand[c];
        const float rel = est[c] - mean;
        const double g2 = gsub2[c] > 1e-30 ? gsub2[c] : 1e-30;
        float f[32];
        f[0] = y;
        f[1] = leta[c];
        f[2] = lg[c];
        f[3] = est[c];
        f[4] = rel;
        f[5] = rel / sd;
        f[6] = (float)log(g2) - lg[c];
        f[7] = kmax;
        f[8] = ksum * is;
        f[9] = sqrtf(ksq * is);
        f[10] = kmin;
        f[11] = c1 * is;
        f[12] = c5 * is;
        f[13] = c9 * is;
        f[14] = log1pf((float)qvcand[c]);
        f[15] = (float)s / (float)(B > 0 ? B : 1);
        f[16] = log1pf((float)s);
        f[17] = mean;
        f[18] = sd;
        f[19] = leta[c] - mleta;
        f[20] = lg[c] - mlg;
        f[21] = rel * rel;
        f[22] = y * rel;
        f[23] = kmax * rel;
        f[24] = expf(fmaxf(est[c] - emax, -20.f));
        f[25] = est[c] - emax;
        f[26] = (emax > emin) ? (est[c] - emin) / (emax - emin) : 0.f;
        f[27] = npos ? (float)npos[0] / (float)s : pos_frac;
        f[28] = ksum * is * y;
        f[29] = (float)fmin(fmax(eta[c], 0.0), 1.0);
        f[30] = 1.0f;
        f[31] = 0.0f;
        for (int q = 0; q < 32; q++) F[(size_t)c * 32 + q] = f[q];
    }
}

__global__ void gather_idx_kernel(const double *__restrict__ src, const int32_t *__restrict__ idx, int m,
                                  double *__restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) dst[j] = src[idx[j]];
}

struct MlpBufs {
    int cap = 0;
    float *gpart = nullptr, *lpart = nullptr;
};
thread_local MlpBufs g_mlp;

void mlp_reserve(int nblk) {
    if (g_mlp.cap >= nblk) return;
    if (g_mlp.gpart) irwls_dev_free(g_mlp.gpart);
    if (g_mlp.lpart) irwls_dev_free(g_mlp.lpart);
    g_mlp.gpart = (float *)irwls_dev_alloc((size_t)nblk * NP * sizeof(float));
    g_mlp.lpart = (float *)irwls_dev_alloc((size_t)nblk * sizeof(float));
    g_mlp.cap = nblk;
}

}  // namespace

extern "C" {

int irwls_mlp_nparams(void) { return NP; }
int irwls_mlp_npacked(void) { return PK_N; }

void irwls_dev_mlp_grad(const float *P, const float *X, const float *T, int nb, float *grad, float *loss) {
    if (nb <= 0) return;
    const int nblk = (nb + SB - 1) / SB;
    mlp_reserve(nblk);
    TimerScope ts(IRWLS_T_MLP, 6.0 * nb * (double)(NIN * NH + NH * NH + NH), 4.0 * NP * nblk);
    ts.launch(mlp_fwd_bwd, dim3(nblk), dim3(256), 0, P, X, T, nb, g_mlp.gpart, g_mlp.lpart);
    ts.launch(mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, g_mlp.gpart, g_mlp.lpart, nblk, nb, grad, loss);
}

void irwls_dev_mlp_adam(float *P, const float *grad, float *m, float *v, int step, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(mlp_adam, dim3((NP + 255) / 256), dim3(256), 0, stream(), P, grad, m, v, lr, b1, b2, eps, c1,
                       c2);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_mlp_forward_f32(const float *P, const float *X, int nb, float *Y) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(mlp_forward_f32, dim3((nb + 63) / 64), dim3(64), 0, stream(), P, X, nb, Y);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_mlp_pack_bf16(const float *P, uint16_t *Q) {
    hipLaunchKernelGGL(mlp_pack, dim3((PK_N + 255) / 256), dim3(256), 0, stream(), P, Q);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_mlp_infer_bf16(const uint16_t *Q, const float *P, const float *X, int nb, float *Y) {
    if (nb <= 0) return;
    TimerScope ts(IRWLS_T_MLP, 2.0 * nb * (double)(NIN * NH + NH * NH + NH), 2.0 * PK_N + 4.0 * nb * NIN);
    ts.launch(mlp_infer_bf16, dim3((nb + SB - 1) / SB), dim3(256), 0, Q, P, X, nb, Y);
}

void irwls_dev_sgma_features(int P, int s, int B, const double *KNC, int ldk, const double *eta, const double *gsub,
                             const double *gsub2, const double *ycand, const double *qvcand, float pos_frac,
                             const int32_t *npos, float *F) {
    if (P <= 0) return;
    if (P > FMAXP) {
        fprintf(stderr, "libirwls: irwls_dev_sgma_features: pool of %d candidates (at most %d)\n", P, FMAXP);
        exit(2);
    }
    float *kst = (float *)scratch(2, (size_t)P * 8 * sizeof(float));
    hipLaunchKernelGGL(knc_stats_kernel, dim3((P + 3) / 4), dim3(256), 0, stream(), P, s, KNC, ldk, kst);
    IRWLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(sgma_features_kernel, dim3(1), dim3(256), 0, stream(), P, s, B, KNC, ldk, eta, gsub, gsub2,
                       ycand, qvcand, pos_frac, npos, kst, F);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_gather_idx(const double *src, const int32_t *idx, int m, double *dst) {
    if (m <= 0) return;
    hipLaunchKernelGGL(gather_idx_kernel, dim3((m + 255) / 256), dim3(256), 0, stream(), src, idx, m, dst);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_release_mlp_buffers(void) {
    if (g_mlp.gpart) irwls_dev_free(g_mlp.gpart);
    if (g_mlp.lpart) irwls_dev_free(g_mlp.lpart);
    g_mlp.gpart = nullptr;
    g_mlp.lpart = nullptr;
    g_mlp.cap = 0;
}

}  // extern "C"
