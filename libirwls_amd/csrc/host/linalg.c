/*
 * Dense linear-algebra entry points of the reference interface
 * (include/ParallelAlgorithms.h:53-200), backed by the device solver.
 */
#include <string.h>
#include "internal.h"
#include "ParallelAlgorithms.h"

/* reference: src/ParallelAlgorithms.c:91-104 (per-thread host scratch).  Here:
   the device context; scratch grows on demand, so `size` needs no action. */
void initMemory(int Threads, int size) {
    (void)Threads;
    (void)size;
    irwls_dev_init();
}

/* reference: src/ParallelAlgorithms.c:118-131 */
void freeMemory(int Threads) {
    (void)Threads;
    irwls_mixed_release();
    irwls_dev_trim();
}

/* reference: src/ParallelAlgorithms.c:147-156 */
void updateMemory(int Threads, int size) {
    (void)Threads;
    (void)size;
}

/* reference: src/ParallelAlgorithms.c:361-447.  The reference factors and
   solves the n x n system at row/column 0 with leading dimension n, ignores
   the LAPACK failure code, leaves the factor in matrix1 with its strict upper
   triangle zeroed, and solves one right-hand side. */
void ParallelLinearSystem(double *matrix1, int r1, int c1, int ro1, int co1, double *matrix2, int r2, int c2, int ro2,
                          int co2, int n, int m, double *result, int rr, int cr, int ror, int cor, int nCores) {
    (void)r1; (void)c1; (void)ro1; (void)co1; (void)r2; (void)c2; (void)ro2; (void)co2;
    (void)m; (void)rr; (void)cr; (void)ror; (void)cor; (void)nCores;
    if (n <= 0) return;
    irwls_dev_init();
    /* device copy with one spare row per column for the right-hand side */
    const int lda = n + 1;
    double *h = (double *)irwls_xcalloc((size_t)lda * n, sizeof(double));
    for (int j = 0; j < n; j++) memcpy(h + (size_t)j * lda, matrix1 + (size_t)j * n, (size_t)n * sizeof(double));
    double *dA = (double *)irwls_dev_alloc((size_t)lda * n * sizeof(double));
    double *db = (double *)irwls_dev_alloc((size_t)n * sizeof(double));
    double *dx = (double *)irwls_dev_alloc((size_t)n * sizeof(double));
    irwls_dev_h2d(dA, h, (size_t)lda * n * sizeof(double));
    irwls_dev_h2d(db, matrix2, (size_t)n * sizeof(double));
    (void)irwls_dev_posv_aug(dA, n, lda, db, dx);
    irwls_dev_d2h(h, dA, (size_t)lda * n * sizeof(double));
    irwls_dev_d2h(result, dx, (size_t)n * sizeof(double));
    for (int j = 0; j < n; j++) {
        memcpy(matrix1 + (size_t)j * n, h + (size_t)j * lda, (size_t)n * sizeof(double));
        for (int k = 0; k < j; k++) matrix1[(size_t)j * n + k] = 0.0;
    }
    free(h);
    irwls_dev_free(dA);
    irwls_dev_free(db);
    irwls_dev_free(dx);
}

static void vector_matrix(double *m1, int size, double *m2, double *result, int trans) {
    if (size <= 0) return;
    irwls_dev_init();
    const size_t nn = (size_t)size * size;
    double *dA = (double *)irwls_dev_alloc(nn * sizeof(double));
    double *dx = (double *)irwls_dev_alloc((size_t)size * sizeof(double));
    double *dy = (double *)irwls_dev_alloc((size_t)size * sizeof(double));
    irwls_dev_h2d(dA, m2, nn * sizeof(double));
    irwls_dev_h2d(dx, m1, (size_t)size * sizeof(double));
    irwls_dev_gemv(dA, size, size, size, trans, dx, dy);
    irwls_dev_d2h(result, dy, (size_t)size * sizeof(double));
    irwls_dev_free(dA);
    irwls_dev_free(dx);
    irwls_dev_free(dy);
}

/* reference: src/ParallelAlgorithms.c:1318-1337, result = m2 m1 */
void ParallelVectorMatrixT(double *m1, int size, double *m2, double *result, int numThreads) {
    (void)numThreads;
    vector_matrix(m1, size, m2, result, 0);
}

/* reference: src/ParallelAlgorithms.c:1352-1371, result = m1^T m2 */
void ParallelVectorMatrix(double *m1, int size, double *m2, double *result, int numThreads) {
    (void)numThreads;
    vector_matrix(m1, size, m2, result, 1);
}
