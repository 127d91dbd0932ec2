/*
 * Learned candidate proposal for SGMA (src/budgeted-train.c:160-216).
 *
 * Plain SGMA draws 64 candidates per step and scores each by its exact gain
 * |KSM_c - KSC z_c|^2 / eta_c, which reads the whole n x s KSC once per step
 * and costs 2 n s flops per candidate.  The learned proposal draws a larger
 * pool, scores it with a small network on cheap features (the gain estimated
 * on a fixed row subsample, the novelty eta_c, the candidate's kernel values
 * to the current centroids, its label), and computes exact gains only for
 * the network's top candidates (irwls_gpu_options.proposal_keep; 0 takes the
 * network's first choice directly).
 *
 * The network is trained during the same run: the first steps are plain
 * SGMA steps that also compute the features of their candidates and log
 * (features, log exact gain - the step's mean log gain).  After the warm-up
 * the network is fitted by full-batch Adam on standardised features (FP32,
 * irwls_dev_mlp_grad / _adam), the standardisation is folded into the first
 * layer, and the weights are packed to BF16 for inference.  If the fit does
 * not lower the loss the run stays plain SGMA.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "internal.h"

#define NF 32

struct irwls_proposal {
    int n, B, m, pmax;
    int32_t *d_sub;        /* subsample rows (m) */
    double *d_KSCs;        /* m x B column-major: KSC on the subsample */
    double *d_KSMs;        /* m x pmax row-major */
    double *d_gs, *d_gs2;  /* subsample squared residuals (pool) */
    double *d_y, *d_qv;    /* candidate labels and |x|^2 (pool) */
    float *d_F;            /* pool features [pmax][32] */
    float *d_P, *d_G, *d_M, *d_V, *d_loss, *d_score;
    uint16_t *d_Q;
    /* training log */
    float *logF;           /* [cap][32] */
    float *logT;
    int nlog, cap;
    int trained;
    double loss0, loss1;
    double *hy, *hqv;
};

/* deterministic stream for the subsample and the weight initialisation
   (libc rand() stays the candidate draws' alone) */
static uint64_t lcg_next(uint64_t *s) {
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return *s >> 11;
}

irwls_proposal *irwls_proposal_begin(int n, int B, int pmax, int log_cap) {
    irwls_proposal *lp = (irwls_proposal *)irwls_xcalloc(1, sizeof(irwls_proposal));
    lp->n = n;
    lp->B = B;
    lp->pmax = pmax;
    lp->m = n < 4096 ? n : 4096;
    /* subsample: m distinct rows by a partial Fisher-Yates */
    int32_t *perm = (int32_t *)irwls_xmalloc((size_t)n * sizeof(int32_t));
    for (int i = 0; i < n; i++) perm[i] = i;
    uint64_t st = 0x9E3779B97F4A7C15ULL;
    for (int i = 0; i < lp->m; i++) {
        const int j = i + (int)(lcg_next(&st) % (uint64_t)(n - i));
        const int32_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    lp->d_sub = (int32_t *)irwls_dev_alloc((size_t)lp->m * sizeof(int32_t));
    irwls_dev_h2d(lp->d_sub, perm, (size_t)lp->m * sizeof(int32_t));
    free(perm);
    const size_t Bz = (size_t)(B > 0 ? B : 1);
    lp->d_KSCs = (double *)irwls_dev_alloc((size_t)lp->m * Bz * sizeof(double));
    lp->d_KSMs = (double *)irwls_dev_alloc((size_t)lp->m * pmax * sizeof(double));
    lp->d_gs = (double *)irwls_dev_alloc((size_t)pmax * sizeof(double));
    lp->d_gs2 = (double *)irwls_dev_alloc((size_t)pmax * sizeof(double));
    lp->d_y = (double *)irwls_dev_alloc((size_t)pmax * sizeof(double));
    lp->d_qv = (double *)irwls_dev_alloc((size_t)pmax * sizeof(double));
    lp->d_F = (float *)irwls_dev_alloc((size_t)pmax * NF * sizeof(float));
    lp->d_score = (float *)irwls_dev_alloc((size_t)pmax * sizeof(float));
    const int np = irwls_mlp_nparams();
    lp->d_P = (float *)irwls_dev_alloc((size_t)np * sizeof(float));
    lp->d_G = (float *)irwls_dev_alloc((size_t)np * sizeof(float));
    lp->d_M = (float *)irwls_dev_alloc((size_t)np * sizeof(float));
    lp->d_V = (float *)irwls_dev_alloc((size_t)np * sizeof(float));
    lp->d_loss = (float *)irwls_dev_alloc(4 * sizeof(float));
    lp->d_Q = (uint16_t *)irwls_dev_alloc((size_t)irwls_mlp_npacked() * sizeof(uint16_t));
    lp->cap = log_cap > 0 ? log_cap : 1;
    lp->logF = (float *)irwls_xmalloc((size_t)lp->cap * NF * sizeof(float));
    lp->logT = (float *)irwls_xmalloc((size_t)lp->cap * sizeof(float));
    lp->hy = (double *)irwls_xmalloc((size_t)pmax * sizeof(double));
    lp->hqv = (double *)irwls_xmalloc((size_t)pmax * sizeof(double));
    return lp;
}

void irwls_proposal_end(irwls_proposal *lp) {
    if (!lp) return;
    void *dev[] = {lp->d_sub, lp->d_KSCs, lp->d_KSMs, lp->d_gs, lp->d_gs2, lp->d_y, lp->d_qv, lp->d_F,
                   lp->d_score, lp->d_P, lp->d_G, lp->d_M, lp->d_V, lp->d_loss, lp->d_Q};
    for (size_t i = 0; i < sizeof(dev) / sizeof(dev[0]); i++) irwls_dev_free(dev[i]);
    free(lp->logF);
    free(lp->logT);
    free(lp->hy);
    free(lp->hqv);
    free(lp);
}

int irwls_proposal_trained(const irwls_proposal *lp) { return lp && lp->trained; }

/* KSC column s (n entries, device) restricted to the subsample. */
void irwls_proposal_add_centroid(irwls_proposal *lp, int s, const double *d_col) {
    irwls_dev_gather_idx(d_col, lp->d_sub, lp->m, lp->d_KSCs + (size_t)s * lp->m);
}

/* KSC column s on the subsample, computed from the centroid (device index
   d_c) when the full column is not at hand. */
void irwls_proposal_add_centroid_rows(irwls_proposal *lp, irwls_devdata *dd, properties props, int s,
                                      const int32_t *d_c) {
    irwls_rows sub = irwls_devdata_rows(dd, lp->d_sub, lp->m);
    irwls_rows cc = irwls_devdata_rows(dd, d_c, 1);
    irwls_dev_kernel_matrix(&sub, &cc, dd->d, props.kernelType, props.Kgamma, lp->d_KSCs + (size_t)s * lp->m, 1);
}

/* Features of a pool of P candidates (device rows d_c, host indices cidx)
   at step s, given their KNC rows (ld B), Z (ld B) and eta (device). */
void irwls_proposal_features(irwls_proposal *lp, irwls_devdata *dd, const svm_dataset *ds, properties props,
                             int s, int P, const int32_t *d_c, const int32_t *cidx, const double *d_KNC,
                             const double *d_Z, const double *d_eta, float pos_frac, const int32_t *d_npos) {
    irwls_rows sub = irwls_devdata_rows(dd, lp->d_sub, lp->m);
    irwls_rows cr = irwls_devdata_rows(dd, d_c, P);
    irwls_dev_kernel_matrix(&sub, &cr, dd->d, props.kernelType, props.Kgamma, lp->d_KSMs, P);
    irwls_dev_sgma_gain_half(lp->m, s, P, lp->d_KSCs, d_Z, lp->B, lp->d_KSMs, P, lp->d_gs, lp->d_gs2);
    for (int i = 0; i < P; i++) {
        lp->hy[i] = ds->y[cidx[i]];
        lp->hqv[i] = ds->quadratic_value[cidx[i]];
    }
    irwls_dev_h2d(lp->d_y, lp->hy, (size_t)P * sizeof(double));
    irwls_dev_h2d(lp->d_qv, lp->hqv, (size_t)P * sizeof(double));
    irwls_dev_sgma_features(P, s, lp->B, d_KNC, lp->B, d_eta, lp->d_gs, lp->d_gs2, lp->d_y, lp->d_qv, pos_frac,
                            d_npos, lp->d_F);
}

/* Log the pool's features with targets log(gain) - mean over the pool. */
void irwls_proposal_log(irwls_proposal *lp, int P, const double *q, const double *eta) {
    if (lp->nlog + P > lp->cap) return;
    irwls_dev_d2h(lp->logF + (size_t)lp->nlog * NF, lp->d_F, (size_t)P * NF * sizeof(float));
    double mean = 0.0;
    for (int i = 0; i < P; i++) {
        const double g = (eta[i] > 0.0) ? q[i] / eta[i] : 0.0;
        const double lg = log(g > 1e-30 ? g : 1e-30);
        lp->logT[lp->nlog + i] = (float)lg;
        mean += lg;
    }
    mean /= (P > 0 ? P : 1);
    for (int i = 0; i < P; i++) lp->logT[lp->nlog + i] -= (float)mean;
    lp->nlog += P;
}

/* Fit the network on the log; returns 1 when it is used from now on. */
int irwls_proposal_train(irwls_proposal *lp, int steps, float lr, irwls_gpu_stats *st) {
    const int N = lp->nlog;
    if (N < 128) return 0;
    /* standardise the features */
    double mu[NF], sd[NF];
    for (int f = 0; f < NF; f++) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < N; i++) {
            const double v = lp->logF[(size_t)i * NF + f];
            a += v;
            b += v * v;
        }
        mu[f] = a / N;
        const double var = b / N - mu[f] * mu[f];
        sd[f] = var > 1e-12 ? sqrt(var) : 1.0;
    }
    float *Xn = (float *)irwls_xmalloc((size_t)N * NF * sizeof(float));
    for (int i = 0; i < N; i++)
        for (int f = 0; f < NF; f++)
            Xn[(size_t)i * NF + f] = (float)((lp->logF[(size_t)i * NF + f] - mu[f]) / sd[f]);
    float *d_X = (float *)irwls_dev_alloc((size_t)N * NF * sizeof(float));
    float *d_T = (float *)irwls_dev_alloc((size_t)N * sizeof(float));
    irwls_dev_h2d(d_X, Xn, (size_t)N * NF * sizeof(float));
    irwls_dev_h2d(d_T, lp->logT, (size_t)N * sizeof(float));
    /* He-uniform initialisation, zero biases */
    const int np = irwls_mlp_nparams();
    float *P = (float *)irwls_xcalloc((size_t)np, sizeof(float));
    uint64_t s = 0xD1B54A32D192ED03ULL;
    const int nin[3] = {32, 128, 128}, nout[3] = {128, 128, 1};
    const int off[3] = {0, 32 * 128 + 128, 32 * 128 + 128 + 128 * 128 + 128};
    for (int l = 0; l < 3; l++) {
        const double a = sqrt(6.0 / nin[l]);
        for (int e = 0; e < nin[l] * nout[l]; e++)
            P[off[l] + e] = (float)(a * (2.0 * (double)(lcg_next(&s) % 1000003) / 1000003.0 - 1.0));
    }
    irwls_dev_h2d(lp->d_P, P, (size_t)np * sizeof(float));
    irwls_dev_memset0(lp->d_M, (size_t)np * sizeof(float));
    irwls_dev_memset0(lp->d_V, (size_t)np * sizeof(float));
    float l0 = 0.0f, l1 = 0.0f;
    for (int k = 1; k <= steps; k++) {
        irwls_dev_mlp_grad(lp->d_P, d_X, d_T, N, lp->d_G, lp->d_loss);
        if (k == 1) irwls_dev_d2h(&l0, lp->d_loss, sizeof(float));
        irwls_dev_mlp_adam(lp->d_P, lp->d_G, lp->d_M, lp->d_V, k, lr);
    }
    irwls_dev_mlp_grad(lp->d_P, d_X, d_T, N, lp->d_G, lp->d_loss);
    irwls_dev_d2h(&l1, lp->d_loss, sizeof(float));
    lp->loss0 = l0;
    lp->loss1 = l1;
    st->mlp_loss0 = l0;
    st->mlp_loss = l1;
    irwls_dev_free(d_X);
    irwls_dev_free(d_T);
    free(Xn);
    if (!(l1 < l0)) {
        free(P);
        return 0;
    }
    /* fold the standardisation into layer 1: W1'[i][j] = W1[i][j]/sd_i,
       b1'[j] = b1[j] - sum_i mu_i W1[i][j]/sd_i */
    irwls_dev_d2h(P, lp->d_P, (size_t)np * sizeof(float));
    for (int j = 0; j < 128; j++) {
        double b = P[32 * 128 + j];
        for (int i = 0; i < NF; i++) {
            const double w = P[i * 128 + j] / sd[i];
            P[i * 128 + j] = (float)w;
            b -= mu[i] * w;
        }
        P[32 * 128 + j] = (float)b;
    }
    irwls_dev_h2d(lp->d_P, P, (size_t)np * sizeof(float));
    irwls_dev_mlp_pack_bf16(lp->d_P, lp->d_Q);
    free(P);
    lp->trained = 1;
    return 1;
}

/* Scores of the pool (device, P floats) from the last features. */
const float *irwls_proposal_score_dev(irwls_proposal *lp, int P) {
    irwls_dev_mlp_infer_bf16(lp->d_Q, lp->d_P, lp->d_F, P, lp->d_score);
    return lp->d_score;
}

/* The same, read back (host, P floats). */
void irwls_proposal_score(irwls_proposal *lp, int P, float *score) {
    irwls_dev_d2h(score, irwls_proposal_score_dev(lp, P), (size_t)P * sizeof(float));
}
