/*
 * Scalar kernel functions, kept as exact CPU functions for callers that
 * evaluate single pairs.  The trainers and the predictor never call them on
 * their hot path: they evaluate kernel tiles on the GPU.
 */
#ifndef KERNELS_
#define KERNELS_

#include "IOStructures.h"

#ifdef __cplusplus
extern "C" {
#endif

/* k(x_index1, x_index2) within one dataset.  Linear: sparse dot.  RBF on a
   dense dataset: exp(-gamma * sum_k (x_k - y_k)^2) over maxdim positions; RBF
   on a sparse dataset: exp(-gamma*(|x|^2 + |y|^2 - 2 x.y)) with the stored
   norms.  reference: include/kernels.h:55, src/kernels.c:63-118 */
double kernelFunction(svm_dataset dataset, int index1, int index2, properties props);

/* k(dataset row, model support vector); RBF always uses the stored norms.
   reference: include/kernels.h:74, src/kernels.c:137-183 */
double kernelTest(svm_dataset dataset, int index1, model mymodel, int index2);

#ifdef __cplusplus
}
#endif

#endif
