/*
 * ORACLE -- test infrastructure only (see irwls_oracle.h).
 *
 * FP64 CPU restatement of the reference LIBIRWLS path.  It follows the
 * reference's control flow and arithmetic order where that order is visible
 * in the reference source; where the reference delegates to BLAS/LAPACK
 * (dgemm, dpotrf, dpotrs, dtrtri/dtrmm/dsyrk inside its recursive solver) the
 * oracle uses textbook loops, so those parts match to rounding, not bitwise.
 * Row-parallel loops use OpenMP exactly where the reference does; every such
 * loop writes disjoint rows, so results do not depend on the thread count.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>
#include "irwls_oracle.h"

static orc_stats g_stats;
static int g_max_outer = 0, g_max_sgma = 0, g_max_par = 0;

const orc_stats *orc_last_stats(void) { return &g_stats; }
void orc_set_limits(int a, int b, int c) { g_max_outer = a; g_max_sgma = b; g_max_par = c; }
void orc_srand(unsigned int seed) { srand(seed); }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Dataset builders: src/IOStructures.c:88-275 (readTrainFile)               */
/* ------------------------------------------------------------------------ */

/* Builds a labelled dataset from per-row (index, value) lists given in CSR
   form with 1-based feature indices. */
static svm_dataset build_labeled(const int64_t *rowptr, const int32_t *idx1,
                                 const double *val, const double *y, int n) {
    svm_dataset ds;
    int maxidx = 0;
    int64_t nnz = rowptr[n];
    for (int64_t k = 0; k < nnz; k++) if (idx1[k] > maxidx) maxidx = idx1[k];
    /* reader: elements = nnz + n terminators + 2*(maxindex+2) for the means */
    int64_t elements = nnz + n + 2 * (int64_t)(maxidx + 2);
    ds.l = n;
    ds.sparse = 0;
    ds.y = (double *)calloc(n + 2, sizeof(double));
    ds.quadratic_value = (double *)calloc(n + 2, sizeof(double));
    ds.x = (svm_sample **)calloc(n + 2, sizeof(svm_sample *));
    ds.features = (svm_sample *)calloc(elements, sizeof(svm_sample));
    double *mpos = (double *)calloc(maxidx + 1, sizeof(double));
    double *mneg = (double *)calloc(maxidx + 1, sizeof(double));
    double npos = 0.0, nneg = 0.0;
    int64_t j = 0;
    int maxdim = 0;
    for (int i = 0; i < n; i++) {
        ds.x[i] = &ds.features[j];
        ds.y[i] = y[i];
        if (y[i] == 1.0) npos += 1; else nneg += 1;
        int dm = 0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++, dm++, j++) {
            ds.features[j].index = idx1[k];
            ds.features[j].value = val[k];
            /* sparse flag: compare with the dm-th entry of the feature array
               (src/IOStructures.c:207-209) */
            if (ds.features[dm].index != ds.features[j].index) ds.sparse = 1;
            if (y[i] == 1.0) mpos[idx1[k]] += val[k]; else mneg[idx1[k]] += val[k];
            ds.quadratic_value[i] += val[k] * val[k];
            if (idx1[k] > maxdim) maxdim = idx1[k];
        }
        ds.features[j++].index = -1;
    }
    /* positive-class mean row l (src/IOStructures.c:238-250) */
    ds.y[n] = 1.0;
    ds.x[n] = &ds.features[j];
    for (int f = 0; f <= maxidx; f++) {
        if (mpos[f] != 0.0) {
            double v = mpos[f] / npos;
            ds.features[j].index = f;
            ds.features[j].value = v;
            ds.quadratic_value[n] += v * v;
            j++;
        }
    }
    ds.features[j++].index = -1;
    /* negative-class mean row l+1; its norm divides by the positive count
       (src/IOStructures.c:253-265) */
    ds.y[n + 1] = -1.0;
    ds.x[n + 1] = &ds.features[j];
    for (int f = 0; f <= maxidx; f++) {
        if (mneg[f] != 0.0) {
            double v = mneg[f] / npos;
            ds.features[j].index = f;
            ds.features[j].value = mneg[f] / nneg;
            ds.quadratic_value[n + 1] += v * v;
            j++;
        }
    }
    ds.features[j++].index = -1;
    ds.maxdim = maxdim;
    free(mpos);
    free(mneg);
    return ds;
}

svm_dataset orc_dataset_dense(const double *X, const double *y, int n, int d) {
    int64_t *rp = (int64_t *)malloc((n + 1) * sizeof(int64_t));
    int32_t *ix = (int32_t *)malloc((size_t)n * d * sizeof(int32_t));
    for (int i = 0; i <= n; i++) rp[i] = (int64_t)i * d;
    for (int64_t k = 0; k < (int64_t)n * d; k++) ix[k] = (int32_t)(k % d) + 1;
    svm_dataset ds = build_labeled(rp, ix, X, y, n);
    free(rp);
    free(ix);
    return ds;
}

svm_dataset orc_dataset_csr(const int64_t *rowptr, const int32_t *col,
                            const double *val, const double *y, int n) {
    int64_t nnz = rowptr[n];
    int32_t *ix = (int32_t *)malloc((nnz > 0 ? nnz : 1) * sizeof(int32_t));
    for (int64_t k = 0; k < nnz; k++) ix[k] = col[k] + 1;
    svm_dataset ds = build_labeled(rowptr, ix, val, y, n);
    free(ix);
    return ds;
}

/* readUnlabeledFile semantics (src/IOStructures.c:510-625): no mean rows. */
svm_dataset orc_dataset_dense_unlabeled(const double *X, int n, int d) {
    svm_dataset ds;
    ds.l = n;
    ds.sparse = 0;
    ds.maxdim = d;
    ds.y = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    ds.quadratic_value = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    ds.x = (svm_sample **)calloc(n > 0 ? n : 1, sizeof(svm_sample *));
    ds.features = (svm_sample *)calloc((size_t)n * (d + 1) + 1, sizeof(svm_sample));
    int64_t j = 0;
    for (int i = 0; i < n; i++) {
        ds.x[i] = &ds.features[j];
        for (int k = 0; k < d; k++, j++) {
            double v = X[(size_t)i * d + k];
            ds.features[j].index = k + 1;
            ds.features[j].value = v;
            ds.quadratic_value[i] += v * v;
        }
        ds.features[j++].index = -1;
    }
    return ds;
}

void orc_free_dataset(svm_dataset ds) {
    free(ds.y);
    free(ds.quadratic_value);
    free(ds.x);
    free(ds.features);
}

/* ------------------------------------------------------------------------ */
/* Kernels: src/kernels.c:63-183                                             */
/* ------------------------------------------------------------------------ */

static double sparse_dot(const svm_sample *a, const svm_sample *b) {
    double s = 0.0;
    while (a->index != -1 && b->index != -1) {
        if (a->index == b->index) { s += a->value * b->value; a++; b++; }
        else if (a->index > b->index) b++;
        else a++;
    }
    return s;
}

/* norm form: |x|^2 + |y|^2 then -2 x_k y_k per matching index */
static double rbf_norm_form(const svm_sample *a, const svm_sample *b, double qa, double qb, double g) {
    double s = qa + qb;
    while (a->index != -1 && b->index != -1) {
        if (a->index == b->index) { s += -2.0 * a->value * b->value; a++; b++; }
        else if (a->index < b->index) a++;
        else b++;
    }
    return exp(-g * s);
}

double orc_kernel(svm_dataset ds, int i, int j, properties props) {
    const svm_sample *a = ds.x[i], *b = ds.x[j];
    if (props.kernelType == 0) return sparse_dot(a, b);
    if (ds.sparse == 0) {
        /* positional walk over maxdim entries (src/kernels.c:90-97) */
        double s = 0.0;
        for (int k = 0; k < ds.maxdim; k++) {
            double t = a[k].value - b[k].value;
            s += t * t;
        }
        return exp(-props.Kgamma * s);
    }
    return rbf_norm_form(a, b, ds.quadratic_value[i], ds.quadratic_value[j], props.Kgamma);
}

double orc_kernel_test(svm_dataset ds, int i, model m, int j) {
    if (m.kernelType == 0) return sparse_dot(ds.x[i], m.x[j]);
    return rbf_norm_form(ds.x[i], m.x[j], ds.quadratic_value[i], m.quadratic_value[j], m.Kgamma);
}

/* ------------------------------------------------------------------------ */
/* Linear algebra (LAPACK potrf/potrs semantics)                             */
/* ------------------------------------------------------------------------ */

/* Unblocked left-looking factorization of an n x n block (LAPACK dpotf2). */
static int potf2(double *A, int n, int lda) {
    for (int j = 0; j < n; j++) {
        double *cj = A + (size_t)j * lda;
        double ajj = cj[j];
        for (int k = 0; k < j; k++) {
            double l = A[(size_t)k * lda + j];
            ajj -= l * l;
        }
        if (!(ajj > 0.0)) {          /* ajj <= 0 or NaN */
            cj[j] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        cj[j] = ajj;
        for (int i = j + 1; i < n; i++) {
            double s = cj[i];
            for (int k = 0; k < j; k++)
                s -= A[(size_t)k * lda + i] * A[(size_t)k * lda + j];
            cj[i] = s / ajj;
        }
    }
    return 0;
}

/* Blocked right-looking factorization (LAPACK dpotrf, lower): the diagonal
   block by potf2, the panel by a triangular solve, the trailing matrix by a
   rank-nb update.  On a non-positive pivot it returns at once, as dpotrf
   does, leaving the Schur value at the pivot. */
int orc_potrf(double *A, int n, int lda) {
    const int nb = 64;
    for (int k0 = 0; k0 < n; k0 += nb) {
        const int kb = (n - k0 < nb) ? n - k0 : nb;
        double *D = A + (size_t)k0 * lda + k0;
        int info = potf2(D, kb, lda);
        if (info) return k0 + info;
        const int r0 = k0 + kb;
        if (r0 >= n) break;
        /* panel rows r0..n-1: solve x L11^T = a */
        #pragma omp parallel for schedule(static)
        for (int i = r0; i < n; i++) {
            for (int j = 0; j < kb; j++) {
                double s = A[(size_t)(k0 + j) * lda + i];
                for (int k = 0; k < j; k++)
                    s -= A[(size_t)(k0 + k) * lda + i] * D[(size_t)k * lda + j];
                A[(size_t)(k0 + j) * lda + i] = s / D[(size_t)j * lda + j];
            }
        }
        /* trailing lower update A22 -= P P^T */
        #pragma omp parallel for schedule(dynamic, 8)
        for (int j = r0; j < n; j++) {
            double *cj = A + (size_t)j * lda;
            for (int k = 0; k < kb; k++) {
                const double *pk = A + (size_t)(k0 + k) * lda;
                const double ajk = pk[j];
                for (int i = j; i < n; i++) cj[i] -= pk[i] * ajk;
            }
        }
    }
    return 0;
}

void orc_potrs(const double *L, int n, int lda, double *b) {
    for (int i = 0; i < n; i++) {           /* L z = b */
        double s = b[i];
        for (int k = 0; k < i; k++) s -= L[(size_t)k * lda + i] * b[k];
        b[i] = s / L[(size_t)i * lda + i];
    }
    for (int i = n - 1; i >= 0; i--) {      /* L^T x = z */
        double s = b[i];
        for (int k = i + 1; k < n; k++) s -= L[(size_t)i * lda + k] * b[k];
        b[i] = s / L[(size_t)i * lda + i];
    }
}

void orc_linear_system(double *A, int n, const double *rhs, double *x) {
    (void)orc_potrf(A, n, n);           /* failure code ignored, as the reference */
    memcpy(x, rhs, n * sizeof(double));
    orc_potrs(A, n, n, x);
    g_stats.solves++;
}

/* ------------------------------------------------------------------------ */
/* Full IRWLS: src/full-train.c                                              */
/* ------------------------------------------------------------------------ */

int *orc_rpermute(int n) {
    int *p = (int *)malloc((n > 0 ? n : 1) * sizeof(int));
    for (int k = 0; k < n; k++) p[k] = k;
    for (int k = n - 1; k > 0; k--) {
        int j = rand() % (k + 1);
        int t = p[j]; p[j] = p[k]; p[k] = t;
    }
    return p;
}

/* Inner-loop continuation rule, src/full-train.c:158, with C's precedence:
   (iter<5) || (bounds violated && iter<1000 && since<5 && dW/nW > 1e-6). */
int orc_inner_continue(int iter, double minb, double maxb, double C, int since, double dW, double nW) {
    return (iter < 5) || ((minb < 0.0 || maxb > C) && iter < 1000 && since < 5 && dW / nW > 1e-6);
}

/* Re-weighting and regrouping after an inner iteration
   (src/full-train.c:264-300), in the reference's order of tests. */
void orc_regroup(int l, const double *e, const double *y, const double *bNew, double C, double *a, int *grp) {
    for (int i = 0; i < l; i++) {
        const double ey = e[i] * y[i];
        if (ey < 0.0) a[i] = 0.0;
        else if (ey < 1.0 / 10000) a[i] = C * 10000.0;
        else a[i] = 1.0 * y[i] * C / e[i];
        if (ey < 0.0 && grp[i] != 2) grp[i] = 2;
        if (grp[i] == 1 && y[i] * bNew[i] >= 0.99 * C && y[i] * bNew[i] <= 1.01 * C) grp[i] = 3;
        if (a[i] == 0.0 && grp[i] == 1) grp[i] = 2;
        if (grp[i] == 2 && a[i] != 0.0) grp[i] = 1;
    }
}

/* Initial weights and groups of a sub-problem (src/full-train.c:131-155). */
void orc_initial_groups(int l, const double *e, const double *y, const double *beta, double C, double *a,
                        int *grp) {
    for (int i = 0; i < l; i++) {
        a[i] = (e[i] * y[i] < 0) ? 0.0 : 1.0 * y[i] * C / e[i];
        if (a[i] == 0) grp[i] = 2;
        else if (beta[i] == y[i] * C) grp[i] = 3;
        else grp[i] = 1;
    }
}

/* Outer-iteration bookkeeping of trainFULL (src/full-train.c:514-533):
   returns 1 when dW/nW < eta (the run ends); tracks the best ratio and the
   iterations since it improved. */
int orc_outer_track(double dW, double nW, double eta, double *bestN, int *since) {
    const int done = (dW / nW < eta);
    if (dW / nW < *bestN) { *bestN = dW / nW; *since = 0; }
    else (*since)++;
    return done;
}

/* Working-set rebuild (src/full-train.c:540-643): one direct pick per
   (class, state) among the KKT violators, state 0 = beta 0, 1 = free,
   2 = at the bound C; other violators and every free sample are candidates
   SC, the rest go to SIN.  Candidates fill the working set in order when
   fewer than W - nSW, else through rpermute (libc rand()).  Returns the
   number of candidates; found (6 ints, [y==1][state]) receives the flags. */
int orc_ws_rebuild(int n, const double *bNew, const double *e, const double *y, double C, int W, int *SW, int *nSWo,
                   int *SIN, int *nSIno, int *SC, int *found) {
    const double thr = 0.001;
    int f[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int nSW = 0, nSIn = 0, nSC = 0;
    for (int i = 0; i < n; i++) {
        const double by = bNew[i] * y[i];
        int state, viol;
        if (by == C) { state = 2; viol = (e[i] * y[i] < -1.0 * thr); }
        else if (bNew[i] == 0.0) { state = 0; viol = (e[i] * y[i] > thr); }
        else if (by != 0.0 && by != C) { state = 1; viol = (fabs(e[i] * y[i]) > thr); }
        else continue;   /* in no set, as in the reference */
        if (viol) {
            int cls = -1;
            if (y[i] == -1) cls = 0;
            else if (y[i] == 1) cls = 1;
            if (cls >= 0 && f[cls][state] == 0) { SW[nSW++] = i; f[cls][state] = 1; }
            else SC[nSC++] = i;
        } else if (state == 1) {
            SC[nSC++] = i;
        } else {
            SIN[nSIn++] = i;
        }
    }
    if (found) for (int c = 0; c < 2; c++) for (int s2 = 0; s2 < 3; s2++) found[3 * c + s2] = f[c][s2];
    const int ncand = nSC;
    if (nSC < W - nSW) {
        for (int i = 0; i < nSC; i++) SW[nSW++] = SC[i];
    } else {
        int *perm = orc_rpermute(nSC);
        const int space = W - nSW;
        for (int i = 0; i < nSC; i++) {
            if (i < space) SW[nSW++] = SC[perm[i]];
            else SIN[nSIn++] = SC[perm[i]];
        }
        free(perm);
    }
    *nSWo = nSW;
    *nSIno = nSIn;
    return ncand;
}

/* IRWLSpar loop rule (src/budgeted-train.c:399). */
int orc_par_continue(int iter, double dW, double nW, int since) {
    return (iter < 500) && (dW / nW > 1e-6) && (since < 5);
}

/* IRWLSpar bookkeeping after iteration `iter` (already incremented;
   src/budgeted-train.c:479-489): M /= 10 when iter > 10 and the ratio grew
   more than 100 times; best-ratio tracking.  Returns 1 when the iterate is
   the new best. */
int orc_par_track(int iter, double ratio, double *M, double *oldnorm, double *bestDW, int *since) {
    if (iter > 10 && ratio > 100 * *oldnorm) *M = *M / 10.0;
    *oldnorm = ratio;
    if (ratio < *bestDW) { *bestDW = ratio; *since = 0; return 1; }
    (*since)++;
    return 0;
}

double *orc_subIRWLS(svm_dataset ds, properties props, double *GIN, double *e, double *beta) {
    const int l = ds.l;
    const double C = props.C;
    double *a = (double *)calloc(l > 0 ? l : 1, sizeof(double));
    int *grp = (int *)calloc(l > 0 ? l : 1, sizeof(int));
    double *bNew = (double *)calloc(l + 1, sizeof(double));
    double *bAux = (double *)calloc(l + 1, sizeof(double));
    double *bBest = (double *)calloc(l + 1, sizeof(double));
    int *S1 = (int *)calloc(l + 1, sizeof(int));
    int *S3 = (int *)calloc(l > 0 ? l : 1, sizeof(int));
    double *H = (double *)calloc((size_t)(l + 1) * (l + 1), sizeof(double));
    double *et = (double *)calloc(l + 1, sizeof(double));
    double *G13 = (double *)calloc(l + 1, sizeof(double));
    double maxb = 0.0, minb = 0.0;
    int iter = 0, since = 0, nS1 = 0, nS3 = 0;
    double dW = 1e9, nW = 1.0, bestDW = 1e9;

    /* initial weights and groups (src/full-train.c:131-155) */
    orc_initial_groups(l, e, ds.y, beta, C, a, grp);
    for (int i = 0; i < l; i++) {
        if (grp[i] == 3) S3[nS3++] = i;
        else if (grp[i] == 1) S1[nS1++] = i;
    }
    S1[nS1] = l;

    /* loop rule of src/full-train.c:158 (C precedence) */
    while (orc_inner_continue(iter, minb, maxb, C, since, dW, nW)) {
        iter++;
        g_stats.inner_iters++;
        const int N = nS1 + 1;
        memset(bAux, 0, N * sizeof(double));
        memset(et, 0, N * sizeof(double));
        memset(H, 0, (size_t)N * N * sizeof(double));
        /* KKT system (src/full-train.c:169-186) */
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nS1; i++) {
            const int si = S1[i];
            H[(size_t)i * N + nS1] = ds.y[si];
            H[(size_t)nS1 * N + i] = ds.y[si];
            et[i] = 1.0 - G13[i] - GIN[si];
            for (int j = 0; j < nS1; j++) {
                const int sj = S1[j];
                H[(size_t)i * N + j] = orc_kernel(ds, si, sj, props) * ds.y[si] * ds.y[sj];
                if (i == j) H[(size_t)i * N + j] += 1.0 / a[si];
            }
        }
        H[(size_t)nS1 * N + nS1] = 0.0;
        et[nS1] = -G13[nS1] - GIN[l];
        orc_linear_system(H, N, et, bAux);

        /* weights, bounds and change (src/full-train.c:205-230) */
        dW = 0.0; nW = 0.0; maxb = 0.0; minb = 0.0;
        memset(bNew, 0, (l + 1) * sizeof(double));
        for (int i = 0; i < nS1; i++) {
            if (bAux[i] > maxb) maxb = bAux[i];
            if (bAux[i] < minb) minb = bAux[i];
            bNew[S1[i]] = bAux[i] * ds.y[S1[i]];
        }
        for (int i = 0; i < nS3; i++) bNew[S3[i]] = C * ds.y[S3[i]];
        bNew[l] = bAux[nS1];
        for (int i = 0; i < l + 1; i++) {
            double t = bNew[i] - beta[i];
            dW += t * t;
            nW += beta[i] * beta[i];
        }
        /* error update over the changed weights (src/full-train.c:236-249) */
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < l; i++) {
            double ei = e[i];
            for (int j = 0; j < l; j++)
                if (bNew[j] != beta[j]) ei = ei - orc_kernel(ds, i, j, props) * (bNew[j] - beta[j]);
            e[i] = ei - (bNew[l] - beta[l]);
        }
        if (dW / nW < bestDW) {
            bestDW = dW / nW;
            since = 0;
            memcpy(bBest, bNew, (l + 1) * sizeof(double));
        } else {
            since++;
        }
        /* re-weighting and regrouping (src/full-train.c:264-300) */
        orc_regroup(l, e, ds.y, bNew, C, a, grp);
        memcpy(beta, bNew, (l + 1) * sizeof(double));
        nS1 = 0; nS3 = 0;
        for (int i = 0; i < l; i++) {
            if (grp[i] == 1) S1[nS1++] = i;
            else if (grp[i] == 3) { S3[nS3] = i; et[nS3] = C; nS3++; }
        }
        S1[nS1] = l;
        /* G13 = K(S1,S3) (C y) (src/full-train.c:313-344) */
        memset(G13, 0, (nS1 + 1) * sizeof(double));
        if (nS3 > 0) {
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < nS1 + 1; i++) {
                double g = 0.0;
                if (i < nS1) {
                    for (int o = 0; o < nS3; o++)
                        g += et[o] * orc_kernel(ds, S1[i], S3[o], props) * ds.y[S1[i]] * ds.y[S3[o]];
                } else {
                    for (int o = 0; o < nS3; o++) g += et[o] * ds.y[S3[o]];
                }
                G13[i] = g;
            }
        }
    }
    free(a); free(grp); free(bAux); free(bNew); free(S3); free(S1);
    free(et); free(G13); free(H);
    return bBest;
}

double *orc_trainFULL(svm_dataset ds, properties props) {
    const int n = ds.l, W = props.MaxSize;
    const double C = props.C;
    double t0 = now_s();
    memset(&g_stats, 0, sizeof(g_stats));
    svm_dataset sub;
    sub.sparse = ds.sparse;
    sub.maxdim = ds.maxdim;
    sub.features = NULL;
    sub.y = (double *)calloc(W, sizeof(double));
    sub.quadratic_value = (double *)calloc(W, sizeof(double));
    sub.x = (svm_sample **)calloc(W, sizeof(svm_sample *));
    double *e = (double *)calloc(n, sizeof(double));
    double *beta = (double *)calloc(n + 1, sizeof(double));
    double *bNew = (double *)calloc(n + 1, sizeof(double));
    int *SW = (int *)calloc(W, sizeof(int));
    int *SIN = (int *)calloc(n, sizeof(int));
    int *SC = (int *)calloc(n, sizeof(int));
    double *GIN = (double *)calloc(W + 1, sizeof(double));
    double *esub = (double *)calloc(W + 1, sizeof(double));
    double *bsub = (double *)calloc(W + 1, sizeof(double));
    int nSW = 0, nSIn = 0;

    /* initial working set: every 10th sample, up to W (src/full-train.c:411-423) */
    for (int i = 0; i < n; i++) {
        if (i % 10 < 1 && nSW < W) SW[nSW++] = i;
        else SIN[nSIn++] = i;
        e[i] = ds.y[i];
    }
    int done = 0, since = 0;
    double bestN = 1e20;
    while (done == 0 && since < 300) {
        g_stats.outer_iters++;
        /* GIN (src/full-train.c:436-456) */
        if (nSIn > 0) {
            memset(GIN, 0, (nSW + 1) * sizeof(double));
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < nSW + 1; i++) {
                double g = 0.0;
                if (i < nSW) {
                    for (int o = 0; o < nSIn; o++)
                        if (bNew[SIN[o]] != 0.0)
                            g += bNew[SIN[o]] * orc_kernel(ds, SW[i], SIN[o], props) * ds.y[SW[i]];
                } else {
                    for (int o = 0; o < nSIn; o++) g += bNew[SIN[o]];
                }
                GIN[i] = g;
            }
        }
        /* sub-dataset (src/full-train.c:463-474) */
        sub.l = nSW;
        for (int i = 0; i < nSW; i++) {
            sub.y[i] = ds.y[SW[i]];
            sub.quadratic_value[i] = ds.quadratic_value[SW[i]];
            sub.x[i] = ds.x[SW[i]];
            bsub[i] = beta[SW[i]];
            esub[i] = e[SW[i]];
        }
        bsub[nSW] = beta[n];
        double *bt = orc_subIRWLS(sub, props, GIN, esub, bsub);
        memcpy(bNew, beta, (n + 1) * sizeof(double));
        for (int i = 0; i < nSW; i++) bNew[SW[i]] = bt[i];
        bNew[n] = bt[nSW];
        /* error update over all samples (src/full-train.c:497-510) */
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            double ei = e[i];
            for (int j = 0; j < nSW; j++)
                ei = ei - orc_kernel(ds, i, SW[j], props) * (bNew[SW[j]] - beta[SW[j]]);
            e[i] = ei - (bNew[n] - beta[n]);
        }
        free(bt);
        double dW = 0.0, nW = 0.0;
        for (int i = 0; i < n + 1; i++) {
            double t = beta[i] - bNew[i];
            dW += t * t;
            nW += beta[i] * beta[i];
        }
        done = orc_outer_track(dW, nW, props.Eta, &bestN, &since);
        memcpy(beta, bNew, (n + 1) * sizeof(double));

        /* working-set rebuild (src/full-train.c:540-643) */
        orc_ws_rebuild(n, bNew, e, ds.y, C, W, SW, &nSW, SIN, &nSIn, SC, NULL);
        if (props.verbose == 1) { printf("."); fflush(stdout); }
        if (g_max_outer > 0 && g_stats.outer_iters >= g_max_outer) break;
    }
    free(e); free(beta); free(GIN); free(esub); free(bsub);
    free(SW); free(SIN); free(SC);
    free(sub.y); free(sub.quadratic_value); free(sub.x);
    if (props.verbose == 1) printf("\n");
    g_stats.t_total = now_s() - t0;
    return bNew;   /* the last iterate, as the reference returns */
}

/* model packing shared by both trainers */
static model pack_model(properties props, svm_dataset ds, const int *rows, int nrows,
                        const double *w, double bias) {
    model m;
    memset(&m, 0, sizeof(m));
    m.Kgamma = props.Kgamma;
    m.bias = bias;
    m.sparse = ds.sparse;
    m.maxdim = ds.maxdim;
    m.kernelType = props.kernelType;
    int nElem = 0;
    for (int r = 0; r < nrows; r++) {
        const svm_sample *s = ds.x[rows[r]];
        while (s->index != -1) { s++; nElem++; }
        nElem++;
    }
    m.nSVs = nrows;
    m.nElem = nElem;
    m.weights = (double *)calloc(nrows > 0 ? nrows : 1, sizeof(double));
    m.quadratic_value = (double *)calloc(nrows > 0 ? nrows : 1, sizeof(double));
    m.x = (svm_sample **)calloc(nrows > 0 ? nrows : 1, sizeof(svm_sample *));
    m.features = (svm_sample *)calloc(nElem > 0 ? nElem : 1, sizeof(svm_sample));
    int fi = 0;
    for (int r = 0; r < nrows; r++) {
        m.quadratic_value[r] = ds.quadratic_value[rows[r]];
        m.weights[r] = w[r];
        m.x[r] = &m.features[fi];
        const svm_sample *s = ds.x[rows[r]];
        while (s->index != -1) {
            m.features[fi].index = s->index;
            m.features[fi].value = s->value;
            s++; fi++;
        }
        m.features[fi].index = -1;
        fi++;
    }
    return m;
}

model orc_fullModel(properties props, svm_dataset ds, double *beta) {
    int *rows = (int *)malloc((ds.l > 0 ? ds.l : 1) * sizeof(int));
    double *w = (double *)malloc((ds.l > 0 ? ds.l : 1) * sizeof(double));
    int k = 0;
    for (int i = 0; i < ds.l; i++)
        if (beta[i] != 0.0) { rows[k] = i; w[k] = beta[i]; k++; }
    model m = pack_model(props, ds, rows, k, w, beta[ds.l]);
    free(rows);
    free(w);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Budgeted: src/budgeted-train.c                                            */
/* ------------------------------------------------------------------------ */

int *orc_randomCentroids(svm_dataset ds, properties props) {
    int *p = (int *)malloc((ds.l > 0 ? ds.l : 1) * sizeof(int));
    for (int i = 0; i < ds.l; i++) p[i] = i;
    for (int i = ds.l - 1; i >= 0; --i) {
        int j = rand() % (i + 1);
        int t = p[i]; p[i] = p[j]; p[j] = t;
    }
    int *c = (int *)malloc((props.size > 0 ? props.size : 1) * sizeof(int));
    for (int i = 0; i < props.size; i++) c[i] = p[i];
    free(p);
    return c;
}

int *orc_SGMA(svm_dataset ds, properties props, int ncand, double *gains_out) {
    const int n = ds.l, B = props.size;
    double t0 = now_s();
    memset(&g_stats, 0, sizeof(g_stats));
    int *cent = (int *)malloc((B > 0 ? B : 1) * sizeof(int));
    int *cidx = (int *)malloc(ncand * sizeof(int));
    double *gain = (double *)malloc(ncand * sizeof(double));
    double *eta = (double *)malloc(ncand * sizeof(double));
    double *KSC = (double *)malloc((size_t)n * B * sizeof(double));   /* column-major n x B */
    double *KNC = (double *)malloc((size_t)ncand * B * sizeof(double));
    double *Z = (double *)malloc((size_t)ncand * B * sizeof(double));
    double *KSM = (double *)malloc((size_t)ncand * n * sizeof(double));
    /* L and L^-1 of K_C + jitter, column-major with leading dimension = size */
    double *L = (double *)calloc((size_t)B * B, sizeof(double));
    double *Li = (double *)calloc((size_t)B * B, sizeof(double));
    double *Lt = (double *)calloc((size_t)B * B, sizeof(double));
    double *Lit = (double *)calloc((size_t)B * B, sizeof(double));
    double *L2 = (double *)calloc(B > 0 ? B : 1, sizeof(double));
    double *IL2 = (double *)calloc(B > 0 ? B : 1, sizeof(double));
    double knc01 = 0.0;
    int best = 0;
    for (int s = 0; s < B; s++) {
        double value;
        if (s > 1) {
            g_stats.sgma_steps++;
            /* candidate draws, serially in candidate order (src/budgeted-train.c:160-167) */
            for (int i = 0; i < ncand; i++) {
                int ix = rand() % n;
                while (ds.y[ix] != ((i % 2) * 2.0 - 1)) ix = rand() % n;
                cidx[i] = ix;
            }
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < ncand; i++) {
                double *ksm = KSM + (size_t)i * n, *knc = KNC + (size_t)i * B, *z = Z + (size_t)i * B;
                for (int r = 0; r < n; r++) ksm[r] = orc_kernel(ds, cidx[i], r, props);
                for (int c = 0; c < s; c++) { knc[c] = orc_kernel(ds, cidx[i], cent[c], props); z[c] = knc[c]; }
                orc_potrs(L, s, s, z);
                double v = 1.0;
                for (int c = 0; c < s; c++) v = v - knc[c] * z[c];
                eta[i] = v;
                /* KSM -= KSC(:, :s) z */
                for (int r = 0; r < n; r++) {
                    double acc = 0.0;
                    for (int c = 0; c < s; c++) acc += KSC[(size_t)c * n + r] * z[c];
                    ksm[r] = ksm[r] - acc;
                }
                double q = 0.0;
                for (int r = 0; r < n; r++) q += ksm[r] * ksm[r];
                gain[i] = (eta[i] > 0.0) ? (1.0 / eta[i]) * q : 0.0;
            }
            value = gain[0];
            best = 0;
            for (int i = 1; i < ncand; i++)
                if (gain[i] > value) { value = gain[i]; best = i; }
            cent[s] = cidx[best];
            if (gains_out) {
                for (int i = 0; i < ncand; i++) {
                    gains_out[((size_t)(s - 2) * ncand + i) * 2] = cidx[i];
                    gains_out[((size_t)(s - 2) * ncand + i) * 2 + 1] = gain[i];
                }
            }
        } else {
            /* centroids 0 and 1 are the class-mean rows (src/budgeted-train.c:219-233) */
            if (s == 0) cent[0] = n;
            else { cent[1] = n + 1; knc01 = orc_kernel(ds, cent[0], cent[1], props); }
            value = 1.0;
            best = 0;
        }
        if (props.verbose == 1) {
            if (s == 0) printf("Best Error Descent %f, Average of positive data is centroid %d\n", value, s);
            if (s == 1) printf("Best Error Descent %f, Average of negative data is centroid %d\n", value, s);
            if (s > 1) printf("Best Error Descent %f, Data with index %d is centroid %d\n", value, cent[s], s);
        }
        #pragma omp parallel for schedule(static)
        for (int r = 0; r < n; r++) KSC[(size_t)s * n + r] = orc_kernel(ds, r, cent[s], props);
        /* grow L and L^-1 by one row (src/budgeted-train.c:247-279) */
        if (s == 0) {
            Lt[0] = sqrt(orc_kernel(ds, cent[0], cent[0], props) + 0.000001);
            Lit[0] = 1.0 / Lt[0];
        } else {
            const double *k = (s == 1) ? &knc01 : KNC + (size_t)best * B;
            for (int r = 0; r < s; r++) {          /* L2 = L^-1 k */
                double acc = 0.0;
                for (int c = 0; c < s; c++) acc += Li[(size_t)c * s + r] * k[c];
                L2[r] = acc;
            }
            double L3 = orc_kernel(ds, cent[s], cent[s], props) + 0.00001;
            for (int r = 0; r < s; r++) L3 = L3 - L2[r] * L2[r];
            L3 = sqrt(L3);
            double IL3 = 1.0 / L3;
            for (int c = 0; c < s; c++) {          /* IL2 = L2^T L^-1 */
                double acc = 0.0;
                for (int r = 0; r < s; r++) acc += L2[r] * Li[(size_t)c * s + r];
                IL2[c] = acc;
            }
            const int s1 = s + 1;
            for (int c = 0; c < s; c++)
                for (int r = 0; r < s; r++) {
                    Lt[(size_t)c * s1 + r] = L[(size_t)c * s + r];
                    Lit[(size_t)c * s1 + r] = Li[(size_t)c * s + r];
                }
            Lt[(size_t)s1 * s1 - 1] = L3;
            Lit[(size_t)s1 * s1 - 1] = IL3;
            for (int c = 0; c < s; c++) {
                Lt[(size_t)c * s1 + s] = L2[c];
                Lit[(size_t)c * s1 + s] = -IL3 * IL2[c];
                Lt[(size_t)s * s1 + c] = 0.0;
                Lit[(size_t)s * s1 + c] = 0.0;
            }
        }
        double *t1 = L, *t2 = Li;
        L = Lt; Li = Lit; Lt = t1; Lit = t2;
        if (g_max_sgma > 0 && s + 1 >= g_max_sgma) {
            for (int r = s + 1; r < B; r++) cent[r] = cent[s];
            break;
        }
    }
    free(cidx); free(gain); free(eta); free(KSC); free(KNC); free(Z); free(KSM);
    free(L); free(Li); free(Lt); free(Lit); free(L2); free(IL2);
    g_stats.t_total = now_s() - t0;
    return cent;
}

double *orc_IRWLSpar(svm_dataset ds, int *idx, properties props) {
    const int n = ds.l, B = props.size;
    const double C = props.C;
    double t0 = now_s();
    int keep_solves = g_stats.solves;
    memset(&g_stats, 0, sizeof(g_stats));
    g_stats.solves = keep_solves;
    double *KC = (double *)calloc((size_t)B * B, sizeof(double));
    double *KSC = (double *)calloc((size_t)n * B, sizeof(double));    /* row-major n x B */
    double *KSCA = (double *)calloc((size_t)n * B, sizeof(double));
    double *Da = (double *)calloc(n, sizeof(double));
    double *Day = (double *)calloc(n, sizeof(double));
    /* KC = K(C,C) + 1e-5 I (src/budgeted-train.c:336-343) */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++) {
            KC[(size_t)i * B + j] = orc_kernel(ds, idx[i], idx[j], props);
            if (i == j) KC[(size_t)i * B + j] += pow(10, -5);
        }
    double M = 10000.0;
    /* first pass: Da = M, Day = M y, KSCA = KSC unscaled (src/budgeted-train.c:348-358) */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        Da[i] = M;
        Day[i] = ds.y[i] * M;
        for (int j = 0; j < B; j++) {
            double k = orc_kernel(ds, i, idx[j], props);
            KSC[(size_t)i * B + j] = k;
            KSCA[(size_t)i * B + j] = k;
        }
    }
    double *K1 = (double *)calloc((size_t)B * B, sizeof(double));
    double *K2 = (double *)calloc(B, sizeof(double));
    double *beta = (double *)calloc(B, sizeof(double));
    double *bNew = (double *)calloc(B, sizeof(double));
    double *bBest = (double *)calloc(B, sizeof(double));
    double *e = (double *)calloc(n, sizeof(double));
    int *ind = (int *)calloc(n, sizeof(int));
    int iter = 0, since = 0, nA = n;
    double dW = 1e9, nW = 1.0, oldnorm = 0.0, bestDW = 1e9;
    while (orc_par_continue(iter, dW, nW, since)) {
        /* K1 = KC + Phi_A^T Phi_A, K2 = Phi_A^T (sqrt(D) y) (src/budgeted-train.c:401-416) */
        memcpy(K1, KC, (size_t)B * B * sizeof(double));
        if (nA > 0) {
            #pragma omp parallel for schedule(static)
            for (int r = 0; r < B; r++) {
                double k2 = 0.0;
                for (int i = 0; i < nA; i++) k2 += KSCA[(size_t)i * B + r] * Day[i];
                K2[r] = k2;
                for (int c = 0; c < B; c++) {
                    double acc = 0.0;
                    for (int i = 0; i < nA; i++) acc += KSCA[(size_t)i * B + r] * KSCA[(size_t)i * B + c];
                    K1[(size_t)c * B + r] += acc;
                }
            }
        } else {
            memset(K2, 0, B * sizeof(double));
        }
        memset(bNew, 0, B * sizeof(double));
        orc_linear_system(K1, B, K2, bNew);
        dW = 0.0; nW = 0.0;
        for (int i = 0; i < B; i++) {
            double t = bNew[i] - beta[i];
            dW += t * t;
            nW += bNew[i] * bNew[i];
            beta[i] = bNew[i];
        }
        /* e = y - KSC beta (src/budgeted-train.c:433-443) */
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int j = 0; j < B; j++) acc += beta[j] * KSC[(size_t)i * B + j];
            e[i] = ds.y[i] - acc;
        }
        /* D = C/(y e) clipped to M, 0 on the wrong side (src/budgeted-train.c:448-457) */
        for (int i = 0; i < n; i++) {
            if (e[i] * ds.y[i] < 0.0) Da[i] = 0.0;
            else Da[i] = 1.0 * C / (ds.y[i] * e[i]);
            if (Da[i] > M) Da[i] = M;
        }
        nA = 0;
        for (int i = 0; i < n; i++) if (Da[i] != 0.0) ind[nA++] = i;
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nA; i++) {
            const double sd = sqrt(Da[ind[i]]);
            for (int j = 0; j < B; j++) KSCA[(size_t)i * B + j] = sd * KSC[(size_t)ind[i] * B + j];
            Day[i] = sd * ds.y[ind[i]];
        }
        ++iter;
        g_stats.irwlspar_iters = iter;
        if (props.verbose == 1) printf("Iteration %d, nSVs %d, ||deltaW||^2/||W||^2=%f\n", iter, nA, dW / nW);
        const double Mold = M;
        if (orc_par_track(iter, dW / nW, &M, &oldnorm, &bestDW, &since)) memcpy(bBest, bNew, B * sizeof(double));
        if (M != Mold) g_stats.m_decays++;
        if (g_max_par > 0 && iter >= g_max_par) break;
    }
    free(KC); free(KSC); free(KSCA); free(Da); free(Day);
    free(K1); free(K2); free(beta); free(bNew); free(e); free(ind);
    g_stats.t_total = now_s() - t0;
    return bBest;
}

model orc_budgetedModel(properties props, svm_dataset ds, int *centroids, double *beta) {
    return pack_model(props, ds, centroids, props.size, beta, 0.0);
}

/* ------------------------------------------------------------------------ */
/* Prediction: src/LIBIRWLS-predict.c:56-131                                 */
/* ------------------------------------------------------------------------ */

double *orc_softTest(svm_dataset ds, model m) {
    double *p = (double *)malloc((ds.l > 0 ? ds.l : 1) * sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < ds.l; i++) {
        double f = m.bias;
        for (int j = 0; j < m.nSVs; j++) f += m.weights[j] * orc_kernel_test(ds, i, m, j);
        p[i] = f;
    }
    return p;
}

double *orc_test(svm_dataset ds, model m) {
    double *p = orc_softTest(ds, m);
    for (int i = 0; i < ds.l; i++) p[i] = (p[i] >= 0.0) ? 1.0 : -1.0;
    return p;
}

/* ------------------------------------------------------------------------ */
/* Model file bytes: src/IOStructures.c:773-788                              */
/* ------------------------------------------------------------------------ */

size_t orc_model_serialize(const model *m, unsigned char *buf) {
    size_t sz = 8 + 8 + 5 * 4 + (size_t)m->nSVs * 16 + (size_t)m->nElem * sizeof(svm_sample);
    if (!buf) return sz;
    unsigned char *p = buf;
    memcpy(p, &m->Kgamma, 8); p += 8;
    memcpy(p, &m->bias, 8); p += 8;
    memcpy(p, &m->maxdim, 4); p += 4;
    memcpy(p, &m->kernelType, 4); p += 4;
    memcpy(p, &m->sparse, 4); p += 4;
    memcpy(p, &m->nSVs, 4); p += 4;
    memcpy(p, &m->nElem, 4); p += 4;
    memcpy(p, m->weights, (size_t)m->nSVs * 8); p += (size_t)m->nSVs * 8;
    memcpy(p, m->quadratic_value, (size_t)m->nSVs * 8); p += (size_t)m->nSVs * 8;
    memcpy(p, m->features, (size_t)m->nElem * sizeof(svm_sample));
    return sz;
}

void orc_free_model(model m) {
    free(m.weights);
    free(m.quadratic_value);
    free(m.x);
    free(m.features);
}
