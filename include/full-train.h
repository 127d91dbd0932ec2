/*
 * Full (working-set) IRWLS SVM training.  Same prototypes and semantics as
 * the reference (include/full-train.h:51-119); the numeric loops run on the
 * GPU through the device layer declared in irwls_device.h.
 */
#ifndef FULLTRAIN_
#define FULLTRAIN_

#include "IOStructures.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fisher-Yates permutation of 0..n-1 drawing rand()%(k+1), k=n-1..1.
   reference: include/full-train.h:51, src/full-train.c:65-77 */
int * rpermute(int n);

/* Inner IRWLS on a working set; returns betaBest (l+1 doubles, last = bias).
   e and beta are updated in place.  reference: include/full-train.h:69,
   src/full-train.c:96-362 */
double* subIRWLS(svm_dataset dataset,properties props, double *GIN, double *e, double *beta);

/* Full training; returns l+1 doubles (signed weights, last = bias).
   reference: include/full-train.h:85, src/full-train.c:373-668 */
double* trainFULL(svm_dataset dataset,properties props);

/* reference: include/full-train.h:94 */
void printFULLInstructions(void) ;

/* reference: include/full-train.h:106 */
properties parseTrainFULLParameters(int* argc, char*** argv);

/* Packs the rows with beta != 0 into a model.
   reference: include/full-train.h:119, src/full-train.c:779-839 */
model calculateFULLModel(properties props, svm_dataset dataset, double * beta );

#ifdef __cplusplus
}
#endif

#endif
