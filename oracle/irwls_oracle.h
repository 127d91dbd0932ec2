/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64, single-host restatement of the reference LIBIRWLS training
 * and prediction path, used as the checker for the GPU library.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library never links or calls it.
 *
 * Each function cites the reference file:line it restates.  How each one is
 * pinned (closed forms, LAPACK via scipy, brute force) is listed in the test
 * files tests/test_oracle_*.py.
 */
#ifndef IRWLS_ORACLE_H
#define IRWLS_ORACLE_H

#include <stdint.h>
#include "IOStructures.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run statistics of the last orc_trainFULL / orc_SGMA / orc_IRWLSpar call. */
typedef struct orc_stats {
    int outer_iters;      /* trainFULL outer iterations */
    int inner_iters;      /* subIRWLS iterations summed over the run */
    int solves;           /* linear systems solved */
    int sgma_steps;       /* SGMA steps with candidates */
    int irwlspar_iters;   /* IRWLSpar iterations */
    double t_total;       /* seconds */
    int m_decays;         /* IRWLSpar: times M was divided by 10 */
} orc_stats;

/* Statistics of the last run. */
const orc_stats *orc_last_stats(void);

/* Stop trainFULL after this many outer iterations / SGMA after this many
   budget slots / IRWLSpar after this many iterations (0 = no limit).  Used
   only to time a bounded CPU sample. */
void orc_set_limits(int max_outer, int max_sgma_size, int max_irwlspar_iters);

/* Dataset builders with the reference readers' semantics
   (src/IOStructures.c:88-275): AoS rows ending in index -1, |x|^2 by a
   running sum of squares, the `sparse` flag rule, and the two appended class
   mean rows (including the negative mean's norm divided by the positive
   count, src/IOStructures.c:259).
   Dense: every row lists indices 1..d (explicit zeros included).
   CSR: 0-based column indices, stored as index col+1; rowptr has n+1 entries. */
svm_dataset orc_dataset_dense(const double *X, const double *y, int n, int d);
svm_dataset orc_dataset_csr(const int64_t *rowptr, const int32_t *col,
                            const double *val, const double *y, int n);
/* Unlabelled variant (no mean rows), as readUnlabeledFile builds. */
svm_dataset orc_dataset_dense_unlabeled(const double *X, int n, int d);
void orc_free_dataset(svm_dataset ds);

/* src/kernels.c:63-118 */
double orc_kernel(svm_dataset ds, int i, int j, properties props);
/* src/kernels.c:137-183 */
double orc_kernel_test(svm_dataset ds, int i, model m, int j);

/* Column-major lower Cholesky, blocked as LAPACK dpotrf (block 64), with its
   failure semantics: at the first pivot a_jj <= 0 (or NaN) the Schur value is
   stored in A(j,j) and j+1 is returned; 0 on success. */
int orc_potrf(double *A, int n, int lda);
/* Solve L L^T x = b in place with the lower factor (LAPACK potrs 'L'). */
void orc_potrs(const double *L, int n, int lda, double *b);
/* ParallelLinearSystem semantics (src/ParallelAlgorithms.c:361-447): factor
   in place, ignore the failure code, solve. */
void orc_linear_system(double *A, int n, const double *rhs, double *x);

/* src/full-train.c:65-77 */
int *orc_rpermute(int n);
/* Control-flow decisions of the trainers, factored out so that tables
   derived from the reference text can check them. */
int orc_inner_continue(int iter, double minb, double maxb, double C, int since, double dW, double nW);  /* :158 */
void orc_regroup(int l, const double *e, const double *y, const double *bNew, double C, double *a,
                 int *grp);                                                                            /* :264-300 */
void orc_initial_groups(int l, const double *e, const double *y, const double *beta, double C, double *a,
                        int *grp);                                                                     /* :131-155 */
int orc_outer_track(double dW, double nW, double eta, double *bestN, int *since);                     /* :514-533 */
int orc_ws_rebuild(int n, const double *bNew, const double *e, const double *y, double C, int W, int *SW,
                   int *nSW, int *SIN, int *nSIn, int *SC, int *found);                                /* :540-643 */
int orc_par_continue(int iter, double dW, double nW, int since);           /* src/budgeted-train.c:399 */
int orc_par_track(int iter, double ratio, double *M, double *oldnorm, double *bestDW, int *since); /* :479-489 */

/* src/full-train.c:96-362 */
double *orc_subIRWLS(svm_dataset ds, properties props, double *GIN, double *e, double *beta);
/* src/full-train.c:373-668 */
double *orc_trainFULL(svm_dataset ds, properties props);
/* src/full-train.c:779-839 */
model orc_fullModel(properties props, svm_dataset ds, double *beta);

/* src/budgeted-train.c:73-100 */
int *orc_randomCentroids(svm_dataset ds, properties props);
/* src/budgeted-train.c:114-309, with `ncand` candidates per step (the
   reference uses 64).  If gains_out is not NULL it receives, per step s >= 2,
   ncand (candidate index, gain) pairs at gains_out[(s-2)*ncand*2 ...]. */
int *orc_SGMA(svm_dataset ds, properties props, int ncand, double *gains_out);
/* src/budgeted-train.c:324-505 */
double *orc_IRWLSpar(svm_dataset ds, int *idx, properties props);
/* src/budgeted-train.c:519-570 */
model orc_budgetedModel(properties props, svm_dataset ds, int *centroids, double *beta);

/* src/LIBIRWLS-predict.c:56-86 (no accuracy print) */
double *orc_softTest(svm_dataset ds, model m);
/* src/LIBIRWLS-predict.c:99-131 (no accuracy print) */
double *orc_test(svm_dataset ds, model m);

/* storeModel byte layout (src/IOStructures.c:773-788).  Returns the byte
   count; writes into buf when buf is not NULL. */
size_t orc_model_serialize(const model *m, unsigned char *buf);
void orc_free_model(model m);

/* libc rand()/srand() of this process, so Python drives the same stream. */
void orc_srand(unsigned int seed);

#ifdef __cplusplus
}
#endif

#endif
