/*
 * Budgeted (SGMA + parametric IRWLS) SVM training.  Same prototypes and
 * semantics as the reference (include/budgeted-train.h:47-106).
 */
#ifndef BUDGETEDTRAIN_
#define BUDGETEDTRAIN_

#include "IOStructures.h"

#ifdef __cplusplus
extern "C" {
#endif

/* First props.size entries of a rand()-driven Fisher-Yates permutation.
   reference: include/budgeted-train.h:47, src/budgeted-train.c:73-100 */
int* randomCentroids(svm_dataset dataset,properties props);

/* Sparse Greedy Matrix Approximation: props.size centroid indices.
   reference: include/budgeted-train.h:58, src/budgeted-train.c:114-309 */
int* SGMA(svm_dataset dataset,properties props);

/* Parametric IRWLS on the centroids; returns props.size weights (betaBest).
   reference: include/budgeted-train.h:70, src/budgeted-train.c:324-505 */
double* IRWLSpar(svm_dataset dataset, int* indexes,properties props);

/* reference: include/budgeted-train.h:80 */
void printBudgetedInstructions(void) ;

/* reference: include/budgeted-train.h:92 */
properties parseTrainParameters(int* argc, char*** argv);

/* reference: include/budgeted-train.h:106, src/budgeted-train.c:519-570 */
model calculateBudgetedModel(properties props, svm_dataset dataset, int *centroids, double * beta );

#ifdef __cplusplus
}
#endif

#endif
