/*
 * Full (working-set) IRWLS training: host control in C, numeric loops on the
 * GPU through irwls_device.h.
 *
 * Control flow, comparisons and the libc rand() consumption order follow the
 * reference (src/full-train.c:65-668) so the working-set trajectory is the
 * same; what moves to the device is every kernel evaluation and the dense
 * solve:
 *   GIN           src/full-train.c:436-456  -> irwls_dev_kernel_matvec
 *   E-update      src/full-train.c:497-510  -> irwls_dev_kernel_matvec
 *   K(S1,S1), H   src/full-train.c:169-186  -> kernel block + irwls_dev_build_kkt
 *   solve         src/full-train.c:192-199  -> irwls_dev_potrf / potrs
 *   inner e       src/full-train.c:236-249  -> irwls_dev_gemv_gather
 *   G13           src/full-train.c:313-344  -> irwls_dev_gemv_gather
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "full-train.h"

/* per thread: the batched trainer runs one problem per host thread */
static __thread irwls_gpu_stats g_stats;

void irwls_gpu_last_stats(irwls_gpu_stats *out) { *out = g_stats; }

void irwls_stats_set(const irwls_gpu_stats *s) { g_stats = *s; }

int *rpermute(int n) {
    int *p = (int *)irwls_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int k = 0; k < n; k++) p[k] = k;
    irwls_rand_block rb;
    irwls_rand_begin(&rb);
    for (int k = n - 1; k > 0; k--) {
        int j = irwls_rand_draw(&rb) % (k + 1);
        int t = p[j];
        p[j] = p[k];
        p[k] = t;
    }
    irwls_rand_end(&rb);
    return p;
}

/* Dense solve of the KKT system on the device, right-hand side in ws->rhs
   (formed on the device), solution into ws->x.  FP64: H (N x N, leading
   dimension N+1) is factored with the right-hand side riding in row N, then
   one backward sweep.  Mixed precision: the closed form of the same result
   from two SPD solves (mixed_kkt.c); FP64 when that declines.  `last` is
   the right-hand side's last entry. */
static void kkt_solve(irwls_ws *ws, int N, double last, const int32_t *S1, long gen) {
    double t0 = irwls_now();
    int done = 0;
    if (irwls_opts()->precision == 1 || irwls_opts()->precision == 2) {
        done = irwls_kkt_solve_mixed(ws->H, N - 1, S1, gen, ws->rhs, ws->ys, last, ws->x, &g_stats);
        if (!done && N - 1 >= 128) g_stats.fp64_fallbacks++;
    }
    if (!done) {
        /* failure code ignored, as the reference */
        (void)irwls_dev_posv_aug(ws->H, N, N + 1, ws->rhs, ws->x);
        g_stats.factorizations++;
    }
    g_stats.solves++;
    g_stats.t_solve += irwls_now() - t0;
}

double *irwls_sub_irwls_dev(irwls_ws *ws, irwls_devdata *dd, const int32_t *d_widx, int l, const double *y,
                            properties props, const double *GIN, double *e, double *beta) {
    const double C = props.C;
    irwls_ws_reserve(ws, l);
    double *a = (double *)irwls_xcalloc(l, sizeof(double));
    int *grp = (int *)irwls_xcalloc(l, sizeof(int));
    double *bNew = (double *)irwls_xcalloc(l + 1, sizeof(double));
    double *bAux = (double *)irwls_xcalloc(l + 1, sizeof(double));
    double *bBest = (double *)irwls_xcalloc(l + 1, sizeof(double));
    int32_t *S1 = (int32_t *)irwls_xcalloc(l + 1, sizeof(int32_t));
    int32_t *S3 = (int32_t *)irwls_xcalloc(l + 1, sizeof(int32_t));
    double *et = (double *)irwls_xcalloc(l + 1, sizeof(double));
    double *hbuf = (double *)irwls_xcalloc(l + 1, sizeof(double));
    double *hbuf2 = (double *)irwls_xcalloc(l + 1, sizeof(double));
    int32_t *cols = (int32_t *)irwls_xcalloc(l + 1, sizeof(int32_t));
    double maxb = 0.0, minb = 0.0;
    int iter = 0, since = 0, nS1 = 0, nS3 = 0;
    double dW = 1e9, nW = 1.0, bestDW = 1e9;
    const long gen = irwls_mixed_new_generation();

    /* the working-set kernel block, evaluated once */
    if (l > 0) irwls_kmat(dd, d_widx, l, d_widx, l, props.kernelType, props.Kgamma, ws->Kw, l);
    irwls_dev_h2d(ws->esub, e, (size_t)l * sizeof(double));
    /* the labels, GIN and the weights on the device, where the right-hand
       sides, the weights after each solve and the error updates are formed */
    irwls_dev_h2d(ws->yl, y, (size_t)l * sizeof(double));
    irwls_dev_h2d(ws->gin, GIN, (size_t)(l + 1) * sizeof(double));
    irwls_dev_h2d(ws->beta, beta, (size_t)(l + 1) * sizeof(double));

    /* initial weights and groups (src/full-train.c:131-155) */
    for (int i = 0; i < l; i++) {
        a[i] = (e[i] * y[i] < 0) ? 0.0 : 1.0 * y[i] * C / e[i];
        if (a[i] == 0) grp[i] = 2;
        else if (beta[i] == y[i] * C) { S3[nS3++] = i; grp[i] = 3; }
        else { S1[nS1++] = i; grp[i] = 1; }
    }
    /* G13 (device, entries 0..nS1-1) and its last entry g (host): zero until
       the first regrouping computes them */
    int use_g13 = 0;
    double g13last = 0.0;
    memcpy(ws->hS3, S3, (size_t)nS3 * sizeof(int32_t));
    irwls_ws_upload(ws, 2);   /* S3 of the current groups */
    const int fp64_solve = !(irwls_opts()->precision == 1 || irwls_opts()->precision == 2);

    /* loop rule of src/full-train.c:158 (C precedence) */
    while ((iter < 5) || ((minb < 0.0 || maxb > C) && iter < 1000 && since < 5 && dW / nW > 1e-6)) {
        iter++;
        g_stats.inner_iters++;
        const int N = nS1 + 1;
        /* KKT system: H from the kernel block, the right-hand side
           1 - G13 - GIN[S1] and -G13_n - GIN_l on the device */
        for (int i = 0; i < nS1; i++) {
            hbuf[i] = y[S1[i]];
            hbuf2[i] = a[S1[i]];
        }
        const double last = -g13last - GIN[l];
        memcpy(ws->hS1, S1, (size_t)N * sizeof(int32_t));
        memcpy(ws->hys, hbuf, (size_t)N * sizeof(double));
        memcpy(ws->has, hbuf2, (size_t)N * sizeof(double));
        irwls_ws_upload(ws, 0);
        int batched = 0;
        if (fp64_solve && irwls_batch_worker()) {
            /* a batched run: the whole iteration's device work joins the
               other problems' pending iterations in one batched round (or
               runs below, when no other problem would join it) */
            const double t0 = irwls_now();
            irwls_inner_work wk = {ws->H, N, N + 1, ws->Kw, l, l, ws->S1, ws->ys, ws->as, ws->G13, use_g13, ws->gin,
                                   last, ws->rhs, ws->x, ws->S3, nS3, ws->yl, C, ws->beta, ws->bnew, ws->cols, ws->w,
                                   ws->nc_cst, ws->esub};
            batched = irwls_batch_inner(&wk);
            if (batched) {
                g_stats.factorizations++;
                g_stats.solves++;
                g_stats.t_solve += irwls_now() - t0;
            }
        }
        if (!batched) {
            irwls_dev_kkt_rhs(ws->S1, nS1, ws->G13, use_g13, ws->gin, last, ws->rhs);
            irwls_dev_build_kkt(ws->Kw, l, ws->S1, nS1, ws->ys, ws->as, ws->H, N + 1);
            kkt_solve(ws, N, last, S1, gen);
            /* weights after the solve and the error update over the changed
               weights (src/full-train.c:205-249), on the device */
            irwls_dev_inner_update(ws->x, ws->S1, nS1, ws->S3, nS3, ws->yl, C, l, ws->beta, ws->bnew, ws->cols,
                                   ws->w, ws->nc_cst);
            irwls_dev_gemv_gather_dn(ws->Kw, l, l, ws->cols, ws->w, ws->nc_cst, 1.0, -1.0, ws->esub);
        }
        /* one read-back of the solution and the errors */
        irwls_dev_d2h2(bAux, ws->x, (size_t)N * sizeof(double), e, ws->esub, (size_t)l * sizeof(double));

        /* weights, bounds and change (src/full-train.c:205-230), the same
           values as the device's */
        dW = 0.0;
        nW = 0.0;
        maxb = 0.0;
        minb = 0.0;
        memset(bNew, 0, (size_t)(l + 1) * sizeof(double));
        for (int i = 0; i < nS1; i++) {
            if (bAux[i] > maxb) maxb = bAux[i];
            if (bAux[i] < minb) minb = bAux[i];
            bNew[S1[i]] = bAux[i] * y[S1[i]];
        }
        for (int i = 0; i < nS3; i++) bNew[S3[i]] = C * y[S3[i]];
        bNew[l] = bAux[nS1];
        for (int i = 0; i < l + 1; i++) {
            double t = bNew[i] - beta[i];
            dW += t * t;
            nW += beta[i] * beta[i];
        }
        if (dW / nW < bestDW) {
            bestDW = dW / nW;
            since = 0;
            memcpy(bBest, bNew, (size_t)(l + 1) * sizeof(double));
        } else {
            since++;
        }
        /* re-weighting and regrouping (src/full-train.c:264-300) */
        for (int i = 0; i < l; i++) {
            const double ey = e[i] * y[i];
            if (ey < 0.0) a[i] = 0.0;
            else if (ey < 1.0 / 10000) a[i] = C * 10000.0;
            else a[i] = 1.0 * y[i] * C / e[i];
            if (ey < 0.0 && grp[i] != 2) grp[i] = 2;
            if (grp[i] == 1 && y[i] * bNew[i] >= 0.99 * C && y[i] * bNew[i] <= 1.01 * C) grp[i] = 3;
            if (a[i] == 0.0 && grp[i] == 1) grp[i] = 2;
            if (grp[i] == 2 && a[i] != 0.0) grp[i] = 1;
            beta[i] = bNew[i];
        }
        beta[l] = bNew[l];
        nS1 = 0;
        nS3 = 0;
        for (int i = 0; i < l; i++) {
            if (grp[i] == 1) S1[nS1++] = i;
            else if (grp[i] == 3) {
                S3[nS3] = i;
                et[nS3] = C;
                nS3++;
            }
        }
        /* G13 = K(S1,S3) (C y) (src/full-train.c:313-344), left on the device
           for the next right-hand side; the new S3 goes up with it */
        use_g13 = 0;
        g13last = 0.0;
        memcpy(ws->hS3, S3, (size_t)nS3 * sizeof(int32_t));
        if (nS3 > 0) {
            double g = 0.0;
            for (int o = 0; o < nS3; o++) {
                hbuf[o] = et[o] * y[S3[o]];
                g += et[o] * y[S3[o]];
            }
            for (int i = 0; i < nS1; i++) hbuf2[i] = y[S1[i]];
            g13last = g;
            if (nS1 > 0) {
                memcpy(ws->hS1g, S1, (size_t)nS1 * sizeof(int32_t));
                memcpy(ws->hwg, hbuf, (size_t)nS3 * sizeof(double));
                memcpy(ws->hrs, hbuf2, (size_t)nS1 * sizeof(double));
                irwls_ws_upload(ws, 2);
                irwls_dev_gemv_gather(ws->Kw, l, ws->S1g, nS1, ws->S3, nS3, ws->wg, 0.0, 1.0, 0.0, ws->rs,
                                      ws->G13);
                use_g13 = 1;
            } else {
                irwls_ws_upload(ws, 2);
            }
        } else {
            irwls_ws_upload(ws, 2);
        }
    }
    if (fp64_solve) irwls_batch_inner_end();
    free(a);
    free(grp);
    free(bAux);
    free(bNew);
    free(S1);
    free(S3);
    free(et);
    free(hbuf);
    free(hbuf2);
    free(cols);
    return bBest;
}

double *subIRWLS(svm_dataset dataset, properties props, double *GIN, double *e, double *beta) {
    const double t_start = irwls_now();
    irwls_dev_init();
    memset(&g_stats, 0, sizeof(g_stats));
    const double wait0 = irwls_dev_wait_seconds(0);
    irwls_devdata *dd = irwls_devdata_acquire(&dataset, 0, -1, 0);
    irwls_ws ws;
    memset(&ws, 0, sizeof(ws));
    double *b = irwls_sub_irwls_dev(&ws, dd, NULL, dataset.l, dataset.y, props, GIN, e, beta);
    g_stats.t_total = irwls_now() - t_start;
    g_stats.t_wait = irwls_dev_wait_seconds(0) - wait0;
    irwls_ws_free(&ws);
    irwls_devdata_release(dd);
    return b;
}

double *trainFULL(svm_dataset dataset, properties props) {
    const double t_start = irwls_now();
    memset(&g_stats, 0, sizeof(g_stats));
    const double wait0 = irwls_dev_wait_seconds(0);
    if (props.verbose == 1) printf("\n");
    irwls_dev_init();
    const int n = dataset.l, W = props.MaxSize;
    const double thr = 0.001, C = props.C;
    const double *y = dataset.y;
    irwls_devdata *dd = irwls_devdata_acquire(&dataset, 0, -1, 0);

    double *e = (double *)irwls_xcalloc(n, sizeof(double));
    double *beta = (double *)irwls_xcalloc(n + 1, sizeof(double));
    double *bNew = (double *)irwls_xcalloc(n + 1, sizeof(double));
    int32_t *SW = (int32_t *)irwls_xcalloc(W > 0 ? W : 1, sizeof(int32_t));
    int *SIN = (int *)irwls_xcalloc(n, sizeof(int));
    int *SC = (int *)irwls_xcalloc(n, sizeof(int));
    double *GIN = (double *)irwls_xcalloc(W + 1, sizeof(double));
    double *esub = (double *)irwls_xcalloc(W + 1, sizeof(double));
    double *bsub = (double *)irwls_xcalloc(W + 1, sizeof(double));
    double *ysub = (double *)irwls_xcalloc(W + 1, sizeof(double));
    int32_t *cols = (int32_t *)irwls_xcalloc(n, sizeof(int32_t));
    double *wv = (double *)irwls_xcalloc(n, sizeof(double));
    /* the nonzero weights of SIN in SIN order (cols/wv of the next GIN),
       collected by the working-set rebuild */
    int32_t *gcols = (int32_t *)irwls_xcalloc(n, sizeof(int32_t));
    double *gw = (double *)irwls_xcalloc(n, sizeof(double));
    int ng = 0;
    int32_t *d_SW = (int32_t *)irwls_dev_alloc((size_t)(W > 0 ? W : 1) * sizeof(int32_t));
    int32_t *d_cols = (int32_t *)irwls_dev_alloc((size_t)n * sizeof(int32_t));
    double *d_w = (double *)irwls_dev_alloc((size_t)n * sizeof(double));
    double *d_e = (double *)irwls_dev_alloc((size_t)n * sizeof(double));
    double *d_GIN = (double *)irwls_dev_alloc((size_t)(W + 1) * sizeof(double));
    double *d_ysw = (double *)irwls_dev_alloc((size_t)(W + 1) * sizeof(double));
    irwls_ws ws;
    memset(&ws, 0, sizeof(ws));
    irwls_ws_reserve(&ws, W);
    int nSW = 0, nSIn = 0, nSC = 0;

    /* initial working set: every 10th sample, up to W (src/full-train.c:411-423) */
    for (int i = 0; i < n; i++) {
        if (i % 10 < 1 && nSW < W) SW[nSW++] = i;
        else SIN[nSIn++] = i;
        e[i] = y[i];
    }
    irwls_dev_h2d(d_e, e, (size_t)n * sizeof(double));
    int done = 0, since = 0;
    double bestN = 1e20;
    while (done == 0 && since < 300) {
        g_stats.outer_iters++;
        irwls_dev_h2d(d_SW, SW, (size_t)nSW * sizeof(int32_t));
        /* GIN = y_SW K(SW, SIN_nz) beta (src/full-train.c:436-456).  The
           reference sums every beta over SIN in SIN order; the zeros leave
           that sum as it is, so the nonzero ones in the same order (gcols,
           gw: collected by the rebuild below) give the same value. */
        if (nSIn > 0) {
            memset(GIN, 0, (size_t)(nSW + 1) * sizeof(double));
            double gsum = 0.0;
            for (int o = 0; o < ng; o++) gsum += gw[o];
            if (ng > 0 && nSW > 0) {
                for (int i = 0; i < nSW; i++) ysub[i] = y[SW[i]];
                irwls_dev_h2d(d_cols, gcols, (size_t)ng * sizeof(int32_t));
                irwls_dev_h2d(d_w, gw, (size_t)ng * sizeof(double));
                irwls_dev_h2d(d_ysw, ysub, (size_t)nSW * sizeof(double));
                irwls_kmv(dd, d_SW, nSW, d_cols, ng, props.kernelType, props.Kgamma, d_w, 0.0, 1.0, 0.0, d_ysw,
                          d_GIN);
                irwls_dev_d2h(GIN, d_GIN, (size_t)nSW * sizeof(double));
            }
            GIN[nSW] = gsum;
        }
        /* working-set sub-problem (src/full-train.c:463-480) */
        for (int i = 0; i < nSW; i++) {
            ysub[i] = y[SW[i]];
            bsub[i] = beta[SW[i]];
            esub[i] = e[SW[i]];
        }
        bsub[nSW] = beta[n];
        double *bt = irwls_sub_irwls_dev(&ws, dd, d_SW, nSW, ysub, props, GIN, esub, bsub);
        memcpy(bNew, beta, (size_t)(n + 1) * sizeof(double));
        for (int i = 0; i < nSW; i++) bNew[SW[i]] = bt[i];
        bNew[n] = bt[nSW];
        free(bt);
        /* E-update over all samples (src/full-train.c:497-510) */
        {
            int nc = 0;
            for (int j = 0; j < nSW; j++) {
                const double dlt = bNew[SW[j]] - beta[SW[j]];
                if (dlt != 0.0) {
                    cols[nc] = SW[j];
                    wv[nc] = dlt;
                    nc++;
                }
            }
            irwls_dev_h2d(d_cols, cols, (size_t)nc * sizeof(int32_t));
            irwls_dev_h2d(d_w, wv, (size_t)nc * sizeof(double));
            irwls_kmv(dd, NULL, n, d_cols, nc, props.kernelType, props.Kgamma, d_w, 1.0, -1.0, -(bNew[n] - beta[n]),
                      NULL, d_e);
            irwls_dev_d2h(e, d_e, (size_t)n * sizeof(double));
        }
        /* The change of the weights (src/full-train.c:519-533) and the
           working-set rebuild (:540-643) in one pass over the samples: the
           rebuild reads only bNew, e and y, and the two sums run over
           i = 0..n in index order, as the reference's own loop forms them. */
        double dW = 0.0, nW = 0.0;
        int found[2][3] = {{0, 0, 0}, {0, 0, 0}};
        nSW = 0;
        nSIn = 0;
        nSC = 0;
        ng = 0;
        for (int i = 0; i < n; i++) {
            const double t = beta[i] - bNew[i];
            dW += t * t;
            nW += beta[i] * beta[i];
            const double by = bNew[i] * y[i];
            int state, viol;
            if (by == C) {
                state = 2;
                viol = (e[i] * y[i] < -1.0 * thr);
            } else if (bNew[i] == 0.0) {
                state = 0;
                viol = (e[i] * y[i] > thr);
            } else if (by != 0.0 && by != C) {
                state = 1;
                viol = (fabs(e[i] * y[i]) > thr);
            } else {
                continue;
            }
            if (viol) {
                const int cls = (y[i] == -1) ? 0 : ((y[i] == 1) ? 1 : -1);
                if (cls >= 0 && found[cls][state] == 0) {
                    SW[nSW++] = i;
                    found[cls][state] = 1;
                } else {
                    SC[nSC++] = i;
                }
            } else if (state == 1) {
                SC[nSC++] = i;
            } else {
                SIN[nSIn++] = i;
                if (bNew[i] != 0.0) {
                    gcols[ng] = i;
                    gw[ng++] = bNew[i];
                }
            }
        }
        {
            const double t = beta[n] - bNew[n];
            dW += t * t;
            nW += beta[n] * beta[n];
        }
        if (dW / nW < props.Eta) done = 1;
        memcpy(beta, bNew, (size_t)(n + 1) * sizeof(double));
        if (dW / nW < bestN) {
            bestN = dW / nW;
            since = 0;
        } else {
            since++;
        }
        if (props.verbose == 1) {
            printf(".");
            fflush(stdout);
        }
        if (nSC < W - nSW) {
            for (int i = 0; i < nSC; i++) SW[nSW++] = SC[i];
        } else {
            int *perm = rpermute(nSC);
            const int space = W - nSW;
            for (int i = 0; i < nSC; i++) {
                if (i < space) {
                    SW[nSW++] = SC[perm[i]];
                } else {
                    const int j = SC[perm[i]];
                    SIN[nSIn++] = j;
                    if (bNew[j] != 0.0) {
                        gcols[ng] = j;
                        gw[ng++] = bNew[j];
                    }
                }
            }
            free(perm);
        }
    }
    irwls_ws_free(&ws);
    irwls_dev_free(d_SW);
    irwls_dev_free(d_cols);
    irwls_dev_free(d_w);
    irwls_dev_free(d_e);
    irwls_dev_free(d_GIN);
    irwls_dev_free(d_ysw);
    irwls_devdata_release(dd);
    free(e);
    free(beta);
    free(SW);
    free(SIN);
    free(SC);
    free(GIN);
    free(esub);
    free(bsub);
    free(ysub);
    free(cols);
    free(wv);
    free(gcols);
    free(gw);
    if (props.verbose == 1) printf("\n");
    g_stats.t_total = irwls_now() - t_start;
    g_stats.t_wait = irwls_dev_wait_seconds(0) - wait0;
    return bNew;   /* the last iterate, as the reference returns (src/full-train.c:666) */
}

/* ---- model packing and command line ------------------------------------- */

model calculateFULLModel(properties props, svm_dataset dataset, double *beta) {
    model m;
    memset(&m, 0, sizeof(m));
    m.Kgamma = props.Kgamma;
    m.bias = beta[dataset.l];
    m.sparse = dataset.sparse;
    m.maxdim = dataset.maxdim;
    m.kernelType = props.kernelType;
    int nElem = 0, nSVs = 0;
    for (int i = 0; i < dataset.l; i++) {
        if (beta[i] != 0.0) {
            nSVs++;
            const svm_sample *s = dataset.x[i];
            while (s->index != -1) {
                s++;
                nElem++;
            }
            nElem++;
        }
    }
    m.nSVs = nSVs;
    m.nElem = nElem;
    m.weights = (double *)irwls_xcalloc(nSVs, sizeof(double));
    m.quadratic_value = (double *)irwls_xcalloc(nSVs, sizeof(double));
    m.x = (svm_sample **)irwls_xcalloc(nSVs, sizeof(svm_sample *));
    m.features = (svm_sample *)irwls_xcalloc(nElem, sizeof(svm_sample));
    int k = 0, f = 0;
    for (int i = 0; i < dataset.l; i++) {
        if (beta[i] == 0.0) continue;
        m.quadratic_value[k] = dataset.quadratic_value[i];
        m.weights[k] = beta[i];
        m.x[k] = &m.features[f];
        const svm_sample *s = dataset.x[i];
        for (; s->index != -1; s++, f++) {
            m.features[f].index = s->index;
            m.features[f].value = s->value;
        }
        m.features[f].index = -1;
        f++;
        k++;
    }
    return m;
}

void printFULLInstructions(void) {
    fprintf(stderr, "full-train: This software train the SVM on the given training set and ");
    fprintf(stderr, "generages a model for futures prediction use.\n\n");
    fprintf(stderr, "Usage: full-train [options] training_set_file model_file\n\n");
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "  -k kernel type: (default 1)\n");
    fprintf(stderr, "       0 -- Linear kernel u'*v\n");
    fprintf(stderr, "       1 -- radial basis function: exp(-gamma*|u-v|^2)\n");
    fprintf(stderr, "  -g gamma: set gamma in radial basis kernel function (default 1)\n");
    fprintf(stderr, "       radial basis K(u,v)= exp(-gamma*|u-v|^2)\n");
    fprintf(stderr, "  -c Cost: set SVM Cost (default 1)\n");
    fprintf(stderr, "  -t Threads: Number of threads (default 1)\n");
    fprintf(stderr, "  -w Working set size: Size of the Least Squares problem in every iteration (default 500)\n");
    fprintf(stderr, "  -e eta: Stop criteria (default 0.001)\n");
    fprintf(stderr, "  -f file format: (default 1)\n");
    fprintf(stderr, "       0 -- CSV format (comma separator)\n");
    fprintf(stderr, "       1 -- libsvm format\n");
    fprintf(stderr, "  -p separator: csv separator character (default \",\" if csv format is selected)\n");
    fprintf(stderr, "  -v verbose: (default 1)\n");
    fprintf(stderr, "       0 -- No screen messages\n");
    fprintf(stderr, "       1 -- Screen messages\n");
}

properties parseTrainFULLParameters(int *argc, char ***argv) {
    properties props;
    memset(&props, 0, sizeof(props));
    props.Kgamma = 1.0;
    props.C = 1.0;
    props.Threads = 1;
    props.MaxSize = 500;
    props.Eta = 0.001;
    props.size = 10;
    props.kernelType = 1;
    props.file = 1;
    props.separator = ",";
    props.verbose = 1;
    int i;
    for (i = 1; i < *argc; ++i) {
        if ((*argv)[i][0] != '-') break;
        if (++i >= *argc) {
            printFULLInstructions();
            exit(1);
        }
        const char *name = &(*argv)[i - 1][1];
        char *val = (*argv)[i];
        if (strcmp(name, "g") == 0) props.Kgamma = atof(val);
        else if (strcmp(name, "c") == 0) props.C = atof(val);
        else if (strcmp(name, "e") == 0) props.Eta = atof(val);
        else if (strcmp(name, "t") == 0) props.Threads = atoi(val);
        else if (strcmp(name, "k") == 0) props.kernelType = atoi(val);
        else if (strcmp(name, "w") == 0) props.MaxSize = atoi(val);
        else if (strcmp(name, "s") == 0) props.size = atoi(val);
        else if (strcmp(name, "f") == 0) props.file = atoi(val);
        else if (strcmp(name, "p") == 0) props.separator = val;
        else if (strcmp(name, "v") == 0) props.verbose = atoi(val);
        else {
            fprintf(stderr, "Unknown parameter %s\n", name);
            printFULLInstructions();
            exit(2);
        }
    }
    for (int j = 1; i + j - 1 < *argc; ++j) (*argv)[j] = (*argv)[i + j - 1];
    *argc -= i - 1;
    return props;
}
