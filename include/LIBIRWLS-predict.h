/*
 * Prediction with a trained model.  Same prototypes and semantics as the
 * reference (include/LIBIRWLS-predict.h:48-82).
 */
#ifndef LIBIRWLSPREDICT_
#define LIBIRWLSPREDICT_

#include "IOStructures.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Labels +1 (decision >= 0) / -1; prints "Accuracy: %f" when Labels==1.
   reference: include/LIBIRWLS-predict.h:48, src/LIBIRWLS-predict.c:99-131 */
double *test(svm_dataset dataset, model mymodel,predictProperties props);

/* Decision values bias + sum_j w_j k(x, sv_j).
   reference: include/LIBIRWLS-predict.h:61, src/LIBIRWLS-predict.c:56-86 */
double *softTest(svm_dataset dataset, model mymodel,predictProperties props);

/* reference: include/LIBIRWLS-predict.h:70 */
void printPredictInstructions(void);

/* reference: include/LIBIRWLS-predict.h:82 */
predictProperties parsePredictParameters(int* argc, char*** argv);

#ifdef __cplusplus
}
#endif

#endif
