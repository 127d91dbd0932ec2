/*
 * Dense linear algebra entry points of the reference that callers use
 * directly (reference: include/ParallelAlgorithms.h:53-200).  Here they are
 * backed by the GPU: initMemory/freeMemory create and release the device
 * context and workspace, ParallelLinearSystem runs the device Cholesky and
 * triangular solves.  The reference's recursive helpers (Chol,
 * TriangleInversion, the *Product family) are an implementation detail of its
 * CPU solver and are not part of this interface.
 */
#ifndef PARALLEL_ALGORITHMS_
#define PARALLEL_ALGORITHMS_

#ifdef __cplusplus
extern "C" {
#endif

/* reference: include/ParallelAlgorithms.h:53, src/ParallelAlgorithms.c:91-104.
   Creates the device context; `size` is the largest system order expected. */
void initMemory(int Threads, int size);

/* reference: include/ParallelAlgorithms.h:66, src/ParallelAlgorithms.c:118-131 */
void freeMemory(int Threads);

/* reference: include/ParallelAlgorithms.h:80, src/ParallelAlgorithms.c:147-156 */
void updateMemory(int Threads, int size);

/* Solves matrix1 * result = matrix2 for one right-hand side by Cholesky
   ('L', column-major, n x n starting at row/column 0).  matrix1 is
   overwritten by the factor.  A non-positive pivot is not an error: as with
   LAPACK potrf, the factorization stops there and leaves the Schur value on
   the diagonal, and the solve proceeds with that factor.
   reference: include/ParallelAlgorithms.h:170, src/ParallelAlgorithms.c:361-447 */
void ParallelLinearSystem(double *matrix1,int r1,int c1, int ro1, int co1,double *matrix2,int r2,int c2, int ro2, int co2,int n, int m,double *result,int rr,int cr, int ror, int cor, int nCores);

/* result = m2 * m1 for a column-major size x size m2.
   reference: include/ParallelAlgorithms.h:185, src/ParallelAlgorithms.c:1318-1337 */
void ParallelVectorMatrixT(double *m1,int size,double *m2,double *result, int numThreads);

/* result = m1^T * m2 for a column-major size x size m2.
   reference: include/ParallelAlgorithms.h:200, src/ParallelAlgorithms.c:1352-1371 */
void ParallelVectorMatrix(double *m1,int size,double *m2,double *result, int numThreads);

#ifdef __cplusplus
}
#endif

#endif
