/*
 * Budgeted training: SGMA basis selection and the parametric IRWLS, host
 * control in C over the device layer.
 *
 * The host keeps the reference's control flow, its serial libc rand()
 * draws (src/budgeted-train.c:84,160-165) and its strict-> argmax
 * (:208-216); the device evaluates every kernel block and the dense algebra:
 *   candidate kernel columns   :180        -> irwls_dev_kernel_matrix
 *   z = K_C^-1 k_C, eta        :182-194    -> two GEMMs with L^-1, irwls_dev_sgma_eta
 *   R = KSM - KSC z, |R|^2     :195-198    -> irwls_dev_sgma_gain (fused)
 *   L, L^-1 growth             :247-279    -> GEMVs with L^-1
 *   K1 = KC + Phi_A^T Phi_A    :401-416    -> irwls_dev_wgram / irwls_dev_wvec
 *   solve                      :420-422    -> irwls_dev_posv_aug
 *   e = y - Phi beta, D        :433-465    -> irwls_dev_resid_rows / irwls_dev_irls_weights
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "budgeted-train.h"

static irwls_gpu_stats g_bstats;


int *randomCentroids(svm_dataset dataset, properties props) {
    int *perm = (int *)irwls_xmalloc((size_t)(dataset.l > 0 ? dataset.l : 1) * sizeof(int));
    for (int i = 0; i < dataset.l; ++i) perm[i] = i;
    irwls_rand_block rb;
    irwls_rand_begin(&rb);
    for (int i = dataset.l - 1; i >= 0; --i) {
        int j = irwls_rand_draw(&rb) % (i + 1);
        int t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    irwls_rand_end(&rb);
    int *c = (int *)irwls_xmalloc((size_t)(props.size > 0 ? props.size : 1) * sizeof(int));
    for (int i = 0; i < props.size; i++) c[i] = perm[i];
    free(perm);
    return c;
}

/* k(x_a, x_b) of two device rows by a 1x1 kernel block, left in d_out */
static void kernel_pair_dev(const irwls_devdata *dd, int32_t *d_pair, int a, int b, properties props, double *d_out) {
    int32_t h[2] = {a, b};
    irwls_dev_h2d(d_pair, h, sizeof(h));
    irwls_rows r = irwls_devdata_rows(dd, d_pair, 1);
    irwls_rows c = irwls_devdata_rows(dd, d_pair + 1, 1);
    irwls_dev_kernel_matrix(&r, &c, dd->d, props.kernelType, props.Kgamma, d_out, 1);
}

/* the same, read back */
static double kernel_pair(const irwls_devdata *dd, int32_t *d_pair, int a, int b, properties props, double *d_out) {
    kernel_pair_dev(dd, d_pair, a, b, props, d_out);
    double v = 0.0;
    irwls_dev_d2h(&v, d_out, sizeof(double));
    return v;
}

/* Exact scoring of nc candidates d_c at step s: KSM(:, c) = k(x_i, cand_c)
   (n x nc row-major), KNC(c, e) = k(cand_c, centroid_e) (ld B),
   Z = L^-T L^-1 KNC^T (ld B), eta_c = 1 - KNC_c . Z_c and
   q_c = |KSM(:, c) - KSC Z_c|^2 (src/budgeted-train.c:176-203). */
/* K(rows, centroids 0..s-1) with leading dimension B: from the kept column
   table of the centroids (CSR data), else built for the call. */
static void kmat_centroids(irwls_devdata *dd, properties props, const irwls_rows *rows, const int32_t *d_cent,
                           int s, irwls_coltab *ctab, double *d_K, int B) {
    if (ctab) {
        if (irwls_dev_coltab_size(ctab) != s) {
            fprintf(stderr, "libirwls: SGMA centroid table holds %d columns at step %d\n",
                    irwls_dev_coltab_size(ctab), s);
            exit(2);
        }
        irwls_dev_kernel_matrix_coltab(rows, ctab, props.kernelType, props.Kgamma, d_K, B);
        return;
    }
    irwls_rows ce = irwls_devdata_rows(dd, d_cent, s);
    irwls_dev_kernel_matrix(rows, &ce, dd->d, props.kernelType, props.Kgamma, d_K, B);
}

static void score_exact(irwls_devdata *dd, properties props, const int32_t *d_rows, int n, int B, int s, int nc,
                        const int32_t *d_c, const int32_t *d_cent, irwls_coltab *ctab, const double *d_Li,
                        const double *d_KSC, double *d_KSM, double *d_KNC, double *d_W, double *d_Z, double *d_eta,
                        double *d_q) {
    /* rows: all n rows (d_rows NULL) or the n rows listed; KSC has leading dimension n */
    irwls_rows all = irwls_devdata_rows(dd, d_rows, n);
    irwls_rows cr = irwls_devdata_rows(dd, d_c, nc);
    irwls_dev_kernel_matrix(&all, &cr, dd->d, props.kernelType, props.Kgamma, d_KSM, nc);
    kmat_centroids(dd, props, &cr, d_cent, s, ctab, d_KNC, B);
    irwls_dev_gemm(0, 0, s, nc, s, 1.0, d_Li, B, d_KNC, B, 0.0, d_W, B);
    irwls_dev_gemm(1, 0, s, nc, s, 1.0, d_Li, B, d_W, B, 0.0, d_Z, B);
    irwls_dev_sgma_eta(s, nc, d_KNC, B, d_Z, B, d_eta);
    irwls_dev_sgma_gain(n, s, nc, d_KSC, d_Z, B, d_KSM, nc, d_q);
}

/* Exact scoring of a learned shortlist: the nc candidates d_c sit at pool
   positions d_pos, whose KNC rows, Z columns and eta the pool step already
   computed (the same products per candidate), so only KSM and the gains are
   new; the shortlist's KNC, Z and eta land in d_KNC2, d_Z2, d_eta2. */
static void score_shortlist(irwls_devdata *dd, properties props, const int32_t *d_rows, int n, int B, int s, int nc,
                            const int32_t *d_c, const int32_t *d_pos, const double *d_KSC, double *d_KSM,
                            const double *d_KNC, const double *d_Z, const double *d_eta, double *d_KNC2,
                            double *d_Z2, double *d_eta2, double *d_q) {
    irwls_rows all = irwls_devdata_rows(dd, d_rows, n);
    irwls_rows cr = irwls_devdata_rows(dd, d_c, nc);
    irwls_dev_kernel_matrix(&all, &cr, dd->d, props.kernelType, props.Kgamma, d_KSM, nc);
    irwls_dev_sgma_gather_shortlist(nc, d_pos, s, B, d_KNC, d_Z, d_eta, d_KNC2, d_Z2, d_eta2);
    irwls_dev_sgma_gain(n, s, nc, d_KSC, d_Z2, B, d_KSM, nc, d_q);
}

static int cmp_i32(const void *a, const void *b) {
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* argmax of q/eta with the first maximum winning (src/budgeted-train.c:208-216) */
static int best_gain(int nc, const double *q, const double *eta, double *value) {
    double bestv = 0.0;
    int best = 0;
    for (int i = 0; i < nc; i++) {
        const double g = (eta[i] > 0.0) ? (1.0 / eta[i]) * q[i] : 0.0;
        if (i == 0 || g > bestv) {
            bestv = g;
            best = i;
        }
    }
    *value = bestv;
    return best;
}

int *SGMA(svm_dataset dataset, properties props) {
    const double t_start = irwls_now();
    memset(&g_bstats, 0, sizeof(g_bstats));
    irwls_dev_init();
    const int n = dataset.l, B = props.size;
    const irwls_gpu_options *opt = irwls_opts();
    const int ncand = opt->sgma_candidates;
    /* learned proposal: pool size, shortlist and warm-up */
    const int learned = (opt->proposal == 1) && B > 2 && n > 0;
    int pool = opt->proposal_pool > 0 ? opt->proposal_pool : ncand;   /* default: plain SGMA's own draws */
    if (pool > 1024) pool = 1024;
    if (pool < ncand) pool = ncand;
    int keep = opt->proposal_keep;
    if (keep > pool) keep = pool;
    int warm = opt->proposal_warmup;
    if (warm <= 0) {
        warm = B / 10;
        if (warm < 20) warm = 20;
        if (warm > 100) warm = 100;
    }
    if (warm > B - 2) warm = B - 2;
    const int cmax = learned ? pool : ncand;
    irwls_devdata *dd = irwls_devdata_acquire(&dataset, 1, -1, 0);
    int *cent = (int *)irwls_xmalloc((size_t)(B > 0 ? B : 1) * sizeof(int));
    int32_t *cidx = (int32_t *)irwls_xmalloc((size_t)cmax * sizeof(int32_t));
    int32_t *cidx2 = (int32_t *)irwls_xmalloc((size_t)cmax * sizeof(int32_t));
    /* eta and q side by side (host and device), read back with one copy */
    double *eq = (double *)irwls_xcalloc(2 * (size_t)cmax, sizeof(double));
    double *eta = eq, *q = eq + cmax;
    float *score = (float *)irwls_xmalloc((size_t)cmax * sizeof(float));
    char *taken = (char *)irwls_xmalloc((size_t)cmax);
    const size_t Bz = (size_t)(B > 0 ? B : 1);
    const int kmax = ncand > keep ? ncand : keep;   /* exact-scored candidates per step */
    /* learned proposal with proposal_rows = m < n: every gain of the run (the
       warm-up's candidates and the shortlists) on a fixed random subset of m
       rows (in ascending order) instead of all n, with KSC kept on those rows
       only (d_KSCe, m x B); the full KSC and KSM are then not needed */
    const int me = (learned && opt->proposal_rows > 0 && opt->proposal_rows < n) ? opt->proposal_rows : 0;
    double *d_KSC = me ? NULL : (double *)irwls_dev_alloc((size_t)n * Bz * sizeof(double));   /* n x B, column-major */
    double *d_L = (double *)irwls_dev_alloc(Bz * Bz * sizeof(double));            /* B x B, column-major */
    double *d_Li = (double *)irwls_dev_alloc(Bz * Bz * sizeof(double));
    double *d_KSM = me ? NULL : (double *)irwls_dev_alloc((size_t)n * (kmax > 0 ? kmax : 1) * sizeof(double));
    double *d_KNC = (double *)irwls_dev_alloc((size_t)cmax * Bz * sizeof(double)); /* cmax x B, row-major */
    double *d_W = (double *)irwls_dev_alloc(Bz * cmax * sizeof(double));          /* B x cmax, column-major */
    double *d_Z = (double *)irwls_dev_alloc(Bz * cmax * sizeof(double));
    double *d_eq = (double *)irwls_dev_alloc(2 * (size_t)cmax * sizeof(double));
    double *d_eta = d_eq, *d_q = d_eq + cmax;
    double *d_lv = (double *)irwls_dev_alloc(2 * (Bz + 1) * sizeof(double));
    double *d_v1 = d_lv, *d_v2 = d_lv + Bz + 1;
    double *d_tmp = (double *)irwls_dev_alloc(64 * sizeof(double));
    int32_t *d_cidx = (int32_t *)irwls_dev_alloc((size_t)cmax * sizeof(int32_t));
    int32_t *d_cent = (int32_t *)irwls_dev_alloc(Bz * sizeof(int32_t));
    int32_t *d_pair = (int32_t *)irwls_dev_alloc(4 * sizeof(int32_t));
    irwls_rows all = irwls_devdata_rows(dd, NULL, n);
    irwls_proposal *lp = learned ? irwls_proposal_begin(n, B, cmax, warm * ncand) : NULL;
    int32_t *d_sube = NULL;
    double *d_KSCe = NULL, *d_KSMe = NULL;
    if (me) {
        int32_t *perm = (int32_t *)irwls_xmalloc((size_t)n * sizeof(int32_t));
        for (int i = 0; i < n; i++) perm[i] = i;
        uint64_t st = 0xD1B54A32D192ED03ULL;
        for (int i = 0; i < me; i++) {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            const int j = i + (int)((st >> 11) % (uint64_t)(n - i));
            const int32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        qsort(perm, (size_t)me, sizeof(int32_t), cmp_i32);
        d_sube = (int32_t *)irwls_dev_alloc((size_t)me * sizeof(int32_t));
        irwls_dev_h2d(d_sube, perm, (size_t)me * sizeof(int32_t));
        free(perm);
        d_KSCe = (double *)irwls_dev_alloc((size_t)me * Bz * sizeof(double));
        d_KSMe = (double *)irwls_dev_alloc((size_t)me * (kmax > 0 ? kmax : 1) * sizeof(double));
    }
    /* the centroids' transposed column block, kept and grown by one column per step (CSR data) */
    irwls_coltab *ctab = (!dd->dense && B > 0) ? irwls_dev_coltab_create(dd->d, B) : NULL;
    int full_col = 1;   /* the step's centroid has its full KSC column */
    const double *knc = NULL;       /* device row of k(new centroid, previous centroids) */
    int npos = 0;

    /* Without progress printing (and unless the network alone picks), each
       step's choice stays on the device: the argmax, the chosen column and
       row, the positive count and the shortlist are device kernels, so the
       host only draws candidates and queues launches, and reads the
       centroids back once at the end.  The printing path reads each choice
       back as it is made.  Both give the same centroids. */
    const int devpick = props.verbose != 1 && !(learned && keep == 0);
    int32_t *d_st = (int32_t *)irwls_dev_alloc(4 * sizeof(int32_t));        /* [0] chosen position, [1] positives */
    int32_t *d_cidx2 = (int32_t *)irwls_dev_alloc((size_t)cmax * sizeof(int32_t));
    double *d_knc = (double *)irwls_dev_alloc(Bz * sizeof(double));
    /* the learned shortlist's pool positions and its KNC rows, Z columns and eta */
    int32_t *pos = (int32_t *)irwls_xmalloc((size_t)cmax * sizeof(int32_t));
    int32_t *d_pos = (int32_t *)irwls_dev_alloc((size_t)cmax * sizeof(int32_t));
    double *d_KNC2 = (double *)irwls_dev_alloc((size_t)cmax * Bz * sizeof(double));
    double *d_Z2 = (double *)irwls_dev_alloc(Bz * cmax * sizeof(double));
    double *d_eta2 = (double *)irwls_dev_alloc((size_t)cmax * sizeof(double));
    const double *d_y = dd->y;   /* labels of the device rows (the class-mean rows included) */

    for (int s = 0; s < B; s++) {
        double value = 0.0;
        if (s > 1) {
            g_bstats.sgma_steps++;
            if (!irwls_proposal_trained(lp)) {
                /* candidate draws in the reference's serial order */
                irwls_rand_block rb;
                irwls_rand_begin(&rb);
                for (int i = 0; i < ncand; i++) {
                    int ix = irwls_rand_draw(&rb) % n;
                    while (dataset.y[ix] != ((i % 2) * 2.0 - 1)) ix = irwls_rand_draw(&rb) % n;
                    cidx[i] = ix;
                }
                irwls_rand_end(&rb);
                irwls_dev_h2d(d_cidx, cidx, (size_t)ncand * sizeof(int32_t));
                if (me)
                    score_exact(dd, props, d_sube, me, B, s, ncand, d_cidx, d_cent, ctab, d_Li, d_KSCe, d_KSMe,
                                d_KNC, d_W, d_Z, d_eta, d_q);
                else
                    score_exact(dd, props, NULL, n, B, s, ncand, d_cidx, d_cent, ctab, d_Li, d_KSC, d_KSM, d_KNC,
                                d_W, d_Z, d_eta, d_q);
                if (lp)   /* warm-up: the candidates' features, logged below with their exact gains */
                    irwls_proposal_features(lp, dd, &dataset, props, s, ncand, d_cidx, cidx, d_KNC, d_Z, d_eta,
                                            (float)npos / (float)s, devpick ? d_st + 1 : NULL);
                if (devpick) {
                    irwls_dev_sgma_pick(ncand, d_q, d_eta, d_cidx, s, n, d_y, d_cent, d_st, d_tmp + 16);
                    if (me) {
                        irwls_dev_gather_col_at(d_KSMe, ncand, d_st, me, d_KSCe + (size_t)s * me);
                        full_col = 0;
                    } else {
                        irwls_dev_gather_col_at(d_KSM, ncand, d_st, n, d_KSC + (size_t)s * n);
                    }
                    irwls_dev_copy_row_at(d_KNC, B, d_st, s, d_knc);
                    knc = d_knc;
                    if (lp) irwls_dev_d2h(eq, d_eq, ((size_t)cmax + ncand) * sizeof(double));
                } else {
                    irwls_dev_d2h(eq, d_eq, ((size_t)cmax + ncand) * sizeof(double));
                    const int best = best_gain(ncand, q, eta, &value);
                    cent[s] = cidx[best];
                    /* the chosen candidate's column is the new KSC column */
                    if (me) {
                        irwls_dev_gather_col(d_KSMe, ncand, best, me, d_KSCe + (size_t)s * me);
                        full_col = 0;
                    } else {
                        irwls_dev_gather_col(d_KSM, ncand, best, n, d_KSC + (size_t)s * n);
                    }
                    knc = d_KNC + (size_t)best * B;
                }
                if (lp) {
                    irwls_proposal_log(lp, ncand, q, eta);
                    if (s - 1 == warm) (void)irwls_proposal_train(lp, 300, 3e-3f, &g_bstats);
                }
            } else {
                /* learned proposal: score a larger pool, then exact gains of the shortlist */
                irwls_rand_block rb;
                irwls_rand_begin(&rb);
                for (int i = 0; i < pool; i++) {
                    int ix = irwls_rand_draw(&rb) % n;
                    while (dataset.y[ix] != ((i % 2) * 2.0 - 1)) ix = irwls_rand_draw(&rb) % n;
                    cidx[i] = ix;
                }
                irwls_rand_end(&rb);
                irwls_dev_h2d(d_cidx, cidx, (size_t)pool * sizeof(int32_t));
                irwls_rows cr = irwls_devdata_rows(dd, d_cidx, pool);
                kmat_centroids(dd, props, &cr, d_cent, s, ctab, d_KNC, B);
                irwls_dev_gemm(0, 0, s, pool, s, 1.0, d_Li, B, d_KNC, B, 0.0, d_W, B);
                irwls_dev_gemm(1, 0, s, pool, s, 1.0, d_Li, B, d_W, B, 0.0, d_Z, B);
                irwls_dev_sgma_eta(s, pool, d_KNC, B, d_Z, B, d_eta);
                irwls_proposal_features(lp, dd, &dataset, props, s, pool, d_cidx, cidx, d_KNC, d_Z, d_eta,
                                        (float)npos / (float)s, devpick ? d_st + 1 : NULL);
                if (devpick) {
                    /* shortlist: the highest scores, ties to the earlier draw; then their exact gains */
                    irwls_dev_sgma_topk_pos(pool, irwls_proposal_score_dev(lp, pool), keep, d_cidx, d_cidx2, d_pos);
                    if (me)
                        score_shortlist(dd, props, d_sube, me, B, s, keep, d_cidx2, d_pos, d_KSCe, d_KSMe, d_KNC, d_Z,
                                        d_eta, d_KNC2, d_Z2, d_eta2, d_q);
                    else
                        score_shortlist(dd, props, NULL, n, B, s, keep, d_cidx2, d_pos, d_KSC, d_KSM, d_KNC, d_Z,
                                        d_eta, d_KNC2, d_Z2, d_eta2, d_q);
                    irwls_dev_sgma_pick(keep, d_q, d_eta2, d_cidx2, s, n, d_y, d_cent, d_st, d_tmp + 16);
                    if (me) {
                        irwls_dev_gather_col_at(d_KSMe, keep, d_st, me, d_KSCe + (size_t)s * me);
                        full_col = 0;
                    } else {
                        irwls_dev_gather_col_at(d_KSM, keep, d_st, n, d_KSC + (size_t)s * n);
                    }
                    irwls_dev_copy_row_at(d_KNC2, B, d_st, s, d_knc);
                    knc = d_knc;
                    g_bstats.learned_steps++;
                } else {
                    irwls_proposal_score(lp, pool, score);
                    /* shortlist: the highest scores, ties to the earlier draw */
                    const int k = keep > 0 ? keep : 1;
                    memset(taken, 0, (size_t)pool);
                    for (int r = 0; r < k; r++) {
                        int bi = -1;
                        for (int i = 0; i < pool; i++)
                            if (!taken[i] && (bi < 0 || score[i] > score[bi])) bi = i;
                        taken[bi] = 1;
                        cidx2[r] = cidx[bi];
                        pos[r] = bi;
                        if (r == 0 && keep == 0) {
                            /* the network's choice without exact scoring */
                            cent[s] = cidx[bi];
                            knc = d_KNC + (size_t)bi * B;
                            value = score[bi];
                        }
                    }
                    if (keep > 0) {
                        irwls_dev_h2d(d_cidx2, cidx2, (size_t)keep * sizeof(int32_t));
                        irwls_dev_h2d(d_pos, pos, (size_t)keep * sizeof(int32_t));
                        if (me)
                            score_shortlist(dd, props, d_sube, me, B, s, keep, d_cidx2, d_pos, d_KSCe, d_KSMe, d_KNC,
                                            d_Z, d_eta, d_KNC2, d_Z2, d_eta2, d_q);
                        else
                            score_shortlist(dd, props, NULL, n, B, s, keep, d_cidx2, d_pos, d_KSC, d_KSM, d_KNC, d_Z,
                                            d_eta, d_KNC2, d_Z2, d_eta2, d_q);
                        irwls_dev_d2h(q, d_q, (size_t)keep * sizeof(double));
                        irwls_dev_d2h(eta, d_eta2, (size_t)keep * sizeof(double));
                        const int best = best_gain(keep, q, eta, &value);
                        cent[s] = cidx2[best];
                        if (me) {
                            irwls_dev_gather_col(d_KSMe, keep, best, me, d_KSCe + (size_t)s * me);
                            full_col = 0;
                        } else {
                            irwls_dev_gather_col(d_KSM, keep, best, n, d_KSC + (size_t)s * n);
                        }
                        knc = d_KNC2 + (size_t)best * B;
                    } else {
                        const int32_t c1 = cent[s];
                        irwls_dev_h2d(d_pair, &c1, sizeof(int32_t));
                        irwls_rows cc = irwls_devdata_rows(dd, d_pair, 1);
                        if (me) {
                            irwls_rows re = irwls_devdata_rows(dd, d_sube, me);
                            irwls_dev_kernel_matrix(&re, &cc, dd->d, props.kernelType, props.Kgamma,
                                                    d_KSCe + (size_t)s * me, 1);
                            full_col = 0;
                        } else {
                            irwls_dev_kernel_matrix(&all, &cc, dd->d, props.kernelType, props.Kgamma,
                                                    d_KSC + (size_t)s * n, 1);
                        }
                    }
                    g_bstats.learned_steps++;
                }
            }
        } else {
            cent[s] = (s == 0) ? n : n + 1;
            value = 1.0;
        }
        if (props.verbose == 1) {
            if (s == 0) printf("Best Error Descent %f, Average of positive data is centroid %d\n", value, s);
            if (s == 1) printf("Best Error Descent %f, Average of negative data is centroid %d\n", value, s);
            if (s > 1) printf("Best Error Descent %f, Data with index %d is centroid %d\n", value, cent[s], s);
        }
        if (s <= 1 || !devpick) {
            if (cent[s] == n || (cent[s] < n && dataset.y[cent[s]] > 0)) npos++;
            int32_t cs = cent[s];
            irwls_dev_h2d(d_cent + s, &cs, sizeof(int32_t));
            if (s == 1 && devpick) irwls_dev_h2d(d_st + 1, &npos, sizeof(int32_t));
        }
        if (ctab) {
            irwls_rows ce = irwls_devdata_rows(dd, d_cent, s + 1);
            irwls_dev_coltab_extend(ctab, &ce);
        }
        if (s <= 1) {
            irwls_rows cc = irwls_devdata_rows(dd, d_cent + s, 1);
            if (me) {
                irwls_rows re = irwls_devdata_rows(dd, d_sube, me);
                irwls_dev_kernel_matrix(&re, &cc, dd->d, props.kernelType, props.Kgamma, d_KSCe + (size_t)s * me, 1);
                full_col = 0;
            } else {
                irwls_dev_kernel_matrix(&all, &cc, dd->d, props.kernelType, props.Kgamma, d_KSC + (size_t)s * n, 1);
            }
        }
        if (full_col) {
            if (lp) irwls_proposal_add_centroid(lp, s, d_KSC + (size_t)s * n);
            if (me) irwls_dev_gather_idx(d_KSC + (size_t)s * n, d_sube, me, d_KSCe + (size_t)s * me);
        } else {
            irwls_proposal_add_centroid_rows(lp, dd, props, s, d_cent + s);
        }
        full_col = 1;
        /* grow L and L^-1 of K_C + jitter by one row (src/budgeted-train.c:247-279),
           on the device: L3 and the new rows never come back to the host */
        if (s == 0) {
            const double kcc = kernel_pair(dd, d_pair, cent[s], cent[s], props, d_tmp);
            const double l0 = sqrt(kcc + 0.000001);
            const double il0 = 1.0 / l0;
            irwls_dev_h2d(d_L, &l0, sizeof(double));
            irwls_dev_h2d(d_Li, &il0, sizeof(double));
        } else {
            if (s == 1) {
                const double k01 = kernel_pair(dd, d_pair, cent[0], cent[1], props, d_tmp);
                irwls_dev_h2d(d_tmp, &k01, sizeof(double));
                knc = d_tmp;
            }
            /* kcc = k(c_s, c_s) from the device index of the new centroid */
            irwls_rows cs1 = irwls_devdata_rows(dd, d_cent + s, 1);
            irwls_dev_kernel_matrix(&cs1, &cs1, dd->d, props.kernelType, props.Kgamma, d_tmp + 8, 1);
            irwls_dev_gemv(d_Li, s, s, B, 0, knc, d_v1);        /* L2 = L^-1 k */
            irwls_dev_gemv(d_Li, s, s, B, 1, d_v1, d_v2);       /* IL2 = L2^T L^-1 */
            irwls_dev_sgma_lrow(s, B, d_tmp + 8, d_v1, d_v2, d_L, d_Li);
        }
    }
    if (devpick && B > 2) irwls_dev_d2h(cent + 2, d_cent + 2, (size_t)(B - 2) * sizeof(int32_t));
    irwls_dev_free(d_st);
    irwls_dev_free(d_cidx2);
    irwls_dev_free(d_knc);
    irwls_dev_free(d_pos);
    irwls_dev_free(d_KNC2);
    irwls_dev_free(d_Z2);
    irwls_dev_free(d_eta2);
    free(pos);
    irwls_dev_coltab_destroy(ctab);
    irwls_dev_free(d_KSC);
    irwls_dev_free(d_L);
    irwls_dev_free(d_Li);
    irwls_dev_free(d_KSM);
    irwls_dev_free(d_KNC);
    irwls_dev_free(d_W);
    irwls_dev_free(d_Z);
    irwls_dev_free(d_eq);
    if (me) {
        irwls_dev_free(d_sube);
        irwls_dev_free(d_KSCe);
        irwls_dev_free(d_KSMe);
    }
    irwls_dev_free(d_lv);
    irwls_dev_free(d_tmp);
    irwls_dev_free(d_cidx);
    irwls_dev_free(d_cent);
    irwls_dev_free(d_pair);
    irwls_devdata_release(dd);
    irwls_proposal_end(lp);
    free(cidx);
    free(cidx2);
    free(score);
    free(taken);
    free(eq);
    g_bstats.t_total = irwls_now() - t_start;
    irwls_stats_set(&g_bstats);
    return cent;
}

double *IRWLSpar(svm_dataset dataset, int *indexes, properties props) {
    const double t_start = irwls_now();
    memset(&g_bstats, 0, sizeof(g_bstats));
    irwls_dev_init();
    const int n = dataset.l, B = props.size;
    const double C = props.C;
    irwls_devdata *dd = irwls_devdata_acquire(&dataset, 1, -1, 0);
    const size_t Bz = (size_t)(B > 0 ? B : 1);
    int32_t *hidx = (int32_t *)irwls_xmalloc(Bz * sizeof(int32_t));
    for (int i = 0; i < B; i++) hidx[i] = indexes[i];
    int32_t *d_idx = (int32_t *)irwls_dev_alloc(Bz * sizeof(int32_t));
    irwls_dev_h2d(d_idx, hidx, (size_t)B * sizeof(int32_t));
    double *d_KC = (double *)irwls_dev_alloc(Bz * Bz * sizeof(double));
    double *d_Phi = (double *)irwls_dev_alloc((size_t)n * Bz * sizeof(double));   /* row-major n x B */
    double *d_K1 = (double *)irwls_dev_alloc((Bz + 1) * Bz * sizeof(double));     /* ld B+1 */
    double *d_K2 = (double *)irwls_dev_alloc((Bz + 1) * sizeof(double));
    double *d_beta = (double *)irwls_dev_alloc((Bz + 1) * sizeof(double));
    double *d_e = (double *)irwls_dev_alloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double *d_D = (double *)irwls_dev_alloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double *d_sw = (double *)irwls_dev_alloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int32_t *d_act = (int32_t *)irwls_dev_alloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));

    /* KC = K(C, C) + 1e-5 I (src/budgeted-train.c:336-343) */
    irwls_rows rc = irwls_devdata_rows(dd, d_idx, B);
    irwls_dev_kernel_matrix(&rc, &rc, dd->d, props.kernelType, props.Kgamma, d_KC, B);
    {
        double *kc = (double *)irwls_xmalloc(Bz * Bz * sizeof(double));
        irwls_dev_d2h(kc, d_KC, Bz * Bz * sizeof(double));
        for (int i = 0; i < B; i++) kc[(size_t)i * B + i] += pow(10, -5);
        irwls_dev_h2d(d_KC, kc, Bz * Bz * sizeof(double));
        free(kc);
    }
    /* Phi = K(X, C), row-major (src/budgeted-train.c:348-358) */
    irwls_rows ra = irwls_devdata_rows(dd, NULL, n);
    irwls_dev_kernel_matrix(&ra, &rc, dd->d, props.kernelType, props.Kgamma, d_Phi, B);

    double M = 10000.0;
    int iter = 0, since = 0;
    double dW = 1e9, nW = 1.0, oldnorm = 0.0, bestDW = 1e9;
    double *beta = (double *)irwls_xcalloc(Bz, sizeof(double));
    double *bNew = (double *)irwls_xcalloc(Bz, sizeof(double));
    double *bBest = (double *)irwls_xcalloc(Bz, sizeof(double));
    int nA = n;
    const int32_t *act = NULL;       /* first pass: all rows, unscaled */
    const double *sw = NULL;
    double ymul = M;
    while (iter < 500 && dW / nW > 1e-6 && since < 5) {
        /* K1 = KC + Phi_A^T Phi_A, K2 = Phi_A^T (sqrt(D) y) */
        irwls_dev_wgram(B, d_Phi, act, sw, nA, d_KC, d_K1, B + 1);
        if (nA > 0) irwls_dev_wvec(B, d_Phi, act, sw, dd->y, ymul, nA, d_K2);
        else irwls_dev_memset0(d_K2, Bz * sizeof(double));
        {
            const double t0 = irwls_now();
            (void)irwls_dev_posv_aug(d_K1, B, B + 1, d_K2, d_beta);
            irwls_dev_d2h(bNew, d_beta, (size_t)B * sizeof(double));
            g_bstats.solves++;
            g_bstats.t_solve += irwls_now() - t0;
        }
        dW = 0.0;
        nW = 0.0;
        for (int i = 0; i < B; i++) {
            const double t = bNew[i] - beta[i];
            dW += t * t;
            nW += bNew[i] * bNew[i];
            beta[i] = bNew[i];
        }
        /* e = y - Phi beta; D, active rows (src/budgeted-train.c:433-465) */
        irwls_dev_resid_rows(n, B, d_Phi, d_beta, dd->y, d_e);
        nA = irwls_dev_irls_weights(n, d_e, dd->y, C, M, d_D, d_act, d_sw);
        act = d_act;
        sw = d_sw;
        ymul = 1.0;
        ++iter;
        g_bstats.irwlspar_iters = iter;
        if (props.verbose == 1) printf("Iteration %d, nSVs %d, ||deltaW||^2/||W||^2=%f\n", iter, nA, dW / nW);
        if (iter > 10 && dW / nW > 100 * oldnorm) {
            M = M / 10.0;
            g_bstats.par_m_decays++;
        }
        oldnorm = dW / nW;
        if (dW / nW < bestDW) {
            bestDW = dW / nW;
            since = 0;
            memcpy(bBest, bNew, (size_t)B * sizeof(double));
        } else {
            since++;
        }
        if (irwls_opts()->par_max_iters > 0 && iter >= irwls_opts()->par_max_iters) break;
    }
    irwls_dev_free(d_idx);
    irwls_dev_free(d_KC);
    irwls_dev_free(d_Phi);
    irwls_dev_free(d_K1);
    irwls_dev_free(d_K2);
    irwls_dev_free(d_beta);
    irwls_dev_free(d_e);
    irwls_dev_free(d_D);
    irwls_dev_free(d_sw);
    irwls_dev_free(d_act);
    irwls_devdata_release(dd);
    free(hidx);
    free(beta);
    free(bNew);
    g_bstats.t_total = irwls_now() - t_start;
    irwls_stats_set(&g_bstats);
    return bBest;
}

model calculateBudgetedModel(properties props, svm_dataset dataset, int *centroids, double *beta) {
    model m;
    memset(&m, 0, sizeof(m));
    m.Kgamma = props.Kgamma;
    m.sparse = dataset.sparse;
    m.maxdim = dataset.maxdim;
    m.nSVs = props.size;
    m.bias = 0.0;
    m.kernelType = props.kernelType;
    int nElem = 0;
    for (int i = 0; i < props.size; i++) {
        const svm_sample *s = dataset.x[centroids[i]];
        while (s->index != -1) {
            s++;
            nElem++;
        }
        nElem++;
    }
    m.nElem = nElem;
    m.weights = (double *)irwls_xcalloc((size_t)props.size, sizeof(double));
    memcpy(m.weights, beta, (size_t)props.size * sizeof(double));
    m.quadratic_value = (double *)irwls_xcalloc((size_t)props.size, sizeof(double));
    m.x = (svm_sample **)irwls_xcalloc((size_t)props.size, sizeof(svm_sample *));
    m.features = (svm_sample *)irwls_xcalloc((size_t)nElem, sizeof(svm_sample));
    int f = 0;
    for (int i = 0; i < props.size; i++) {
        m.quadratic_value[i] = dataset.quadratic_value[centroids[i]];
        m.x[i] = &m.features[f];
        const svm_sample *s = dataset.x[centroids[i]];
        for (; s->index != -1; s++, f++) {
            m.features[f].index = s->index;
            m.features[f].value = s->value;
        }
        m.features[f].index = -1;
        f++;
    }
    return m;
}

properties parseTrainParameters(int *argc, char ***argv) {
    properties props;
    memset(&props, 0, sizeof(props));
    props.Kgamma = 1.0;
    props.C = 1.0;
    props.Threads = 1;
    props.MaxSize = 500;
    props.Eta = 0.001;
    props.size = 10;
    props.algorithm = 1;
    props.kernelType = 1;
    props.file = 1;
    props.separator = ",";
    props.verbose = 1;
    int i;
    for (i = 1; i < *argc; ++i) {
        if ((*argv)[i][0] != '-') break;
        if (++i >= *argc) {
            printBudgetedInstructions();
            exit(1);
        }
        const char *name = &(*argv)[i - 1][1];
        char *val = (*argv)[i];
        if (strcmp(name, "g") == 0) props.Kgamma = atof(val);
        else if (strcmp(name, "c") == 0) props.C = atof(val);
        else if (strcmp(name, "e") == 0) props.Eta = atof(val);
        else if (strcmp(name, "k") == 0) props.kernelType = atoi(val);
        else if (strcmp(name, "t") == 0) props.Threads = atoi(val);
        else if (strcmp(name, "w") == 0) props.MaxSize = atoi(val);
        else if (strcmp(name, "s") == 0) props.size = atoi(val);
        else if (strcmp(name, "a") == 0) props.algorithm = atoi(val);
        else if (strcmp(name, "f") == 0) props.file = atoi(val);
        else if (strcmp(name, "p") == 0) props.separator = val;
        else if (strcmp(name, "v") == 0) props.verbose = atoi(val);
        else {
            fprintf(stderr, "Unknown parameter %s\n", name);
            printBudgetedInstructions();
            exit(2);
        }
    }
    for (int j = 1; i + j - 1 < *argc; ++j) (*argv)[j] = (*argv)[i + j - 1];
    *argc -= i - 1;
    return props;
}

void printBudgetedInstructions(void) {
    fprintf(stderr, "budgeted-train: This software train the sparse SVM on the given training set ");
    fprintf(stderr, "and generages a model for futures prediction use.\n\n");
    fprintf(stderr, "Usage: budgeted-train [options] training_set_file model_file\n\n");
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "  -k kernel type: (default 1)\n");
    fprintf(stderr, "       0 -- Linear kernel u'*v\n");
    fprintf(stderr, "       1 -- radial basis function: exp(-gamma*|u-v|^2)\n");
    fprintf(stderr, "  -g gamma: set gamma in radial basis kernel function (default 1)\n");
    fprintf(stderr, "       radial basis K(u,v)= exp(-gamma*|u-v|^2)\n");
    fprintf(stderr, "  -c Cost: set SVM Cost (default 1)\n");
    fprintf(stderr, "  -t Threads: Number of threads (default 1)\n");
    fprintf(stderr, "  -s Classifier size: Size of the classifier (default 1)\n");
    fprintf(stderr, "  -a Algorithm: Algorithm for centroids selection (default 1)\n");
    fprintf(stderr, "       0 -- Random Selection\n");
    fprintf(stderr, "       1 -- SGMA (Sparse Greedy Matrix Approximation)\n");
    fprintf(stderr, "  -f file format: (default 1)\n");
    fprintf(stderr, "       0 -- CSV format (comma separator)\n");
    fprintf(stderr, "       1 -- libsvm format\n");
    fprintf(stderr, "  -p separator: csv separator character (default \",\" if csv format is selected)\n");
    fprintf(stderr, "  -v verbose: (default 1)\n");
    fprintf(stderr, "       0 -- No screen messages\n");
    fprintf(stderr, "       1 -- Screen messages\n");
}
