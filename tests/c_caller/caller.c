/*
 * A C program written against the library's public headers only, the way a
 * caller of the reference API links it: parse the command-line options with
 * the exported parsers, read a libsvm training file, train (full or
 * budgeted), store the model file, read it back, predict a test file and
 * write the labels.  Used by the tests as a compiled caller of the C ABI.
 *
 *   caller full <train> <test> <model> <labels> [trainer flags...]
 *   caller budgeted <train> <test> <model> <labels> [trainer flags...]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "IOStructures.h"
#include "LIBIRWLS-predict.h"
#include "ParallelAlgorithms.h"
#include "budgeted-train.h"
#include "full-train.h"

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: caller full|budgeted train test model labels [flags]\n");
        return 1;
    }
    const int budgeted = strcmp(argv[1], "budgeted") == 0;
    char *train = argv[2], *testf = argv[3], *modelf = argv[4], *labelf = argv[5];
    /* the parsers read the flags and leave the positional arguments */
    int pargc = argc - 5;
    char **pargv = (char **)malloc(sizeof(char *) * (size_t)(pargc + 1));
    pargv[0] = argv[0];
    for (int i = 6; i < argc; i++) pargv[i - 5] = argv[i];
    pargv[pargc] = NULL;
    properties props = budgeted ? parseTrainParameters(&pargc, &pargv) : parseTrainFULLParameters(&pargc, &pargv);

    svm_dataset ds = readTrainFile(train);
    srand(0);
    initMemory(props.Threads, budgeted ? props.size : props.MaxSize + 1);
    model m;
    if (budgeted) {
        int *cent = SGMA(ds, props);
        double *beta = IRWLSpar(ds, cent, props);
        m = calculateBudgetedModel(props, ds, cent, beta);
        free(cent);
        free(beta);
    } else {
        double *beta = trainFULL(ds, props);
        m = calculateFULLModel(props, ds, beta);
        free(beta);
    }
    freeMemory(props.Threads);
    FILE *f = fopen(modelf, "wb");
    if (!f) return 3;
    storeModel(&m, f);
    fclose(f);
    freeModel(m);
    freeDataset(ds);

    model back;
    f = fopen(modelf, "rb");
    if (!f) return 3;
    readModel(&back, f);
    fclose(f);
    svm_dataset td = readTrainFile(testf);
    predictProperties pp;
    memset(&pp, 0, sizeof(pp));
    pp.Labels = 1;
    pp.Threads = 1;
    double *pred = test(td, back, pp);
    writeOutput(labelf, pred, td.l);
    printf("caller: %s, %d training rows, %d support vectors\n", budgeted ? "budgeted" : "full", ds.l, back.nSVs);
    free(pred);
    freeModel(back);
    freeDataset(td);
    free(pargv);
    return 0;
}
