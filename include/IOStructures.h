/*
 * Data model and I/O of the IRWLS SVM trainers: the struct layouts every
 * caller compiles against, the dataset readers and the binary model file.
 *
 * Drop-in boundary.  The five structs below keep the reference's member order
 * and types exactly (reference: include/IOStructures.h:44-122), so existing
 * callers and existing model files keep working.  On x86-64 the sizes are
 * svm_sample 16, svm_dataset 48, properties 72, predictProperties 32,
 * model 72 bytes; libirwls_amd/csrc/host/io.c static-asserts them.
 *
 * Every array a function returns is malloc/calloc-ed; the caller frees it.
 */
#ifndef IOSTRUCTURES_
#define IOSTRUCTURES_

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Training parameters (reference: include/IOStructures.h:44-56). */
typedef struct properties{
    double Kgamma;      /* RBF gamma */
    int kernelType;     /* 0 linear, 1 RBF */
    double C;           /* SVM cost */
    int Threads;        /* host threads (sizes host loops only here) */
    int MaxSize;        /* working-set size W of full training */
    int size;           /* budget B of budgeted training */
    int algorithm;      /* budgeted centroid selection: 0 random, 1 SGMA */
    double Eta;         /* stopping threshold of full training */
    int file;           /* 1 libsvm, 0 CSV */
    char *separator;    /* CSV separator */
    int verbose;        /* 1 prints progress */
}properties;

/* Prediction parameters (reference: include/IOStructures.h:65-72). */
typedef struct predictProperties{
    int Labels;
    int Threads;
    int Soft;
    int file;
    char *separator;
    int verbose;
}predictProperties;

/* Trained model (reference: include/IOStructures.h:81-93). */
typedef struct model{
    double Kgamma;
    int kernelType;
    int sparse;
    int nSVs;
    int nElem;
    double *weights;
    struct svm_sample **x;
    double *quadratic_value;
    int maxdim;
    double bias;
    struct svm_sample* features;
}model;

/* One (index, value) feature; rows end with index -1
   (reference: include/IOStructures.h:102-105). */
typedef struct svm_sample{
    int index;
    double value;
}svm_sample;

/* A labelled dataset.  Readers append two extra rows l and l+1 holding the
   positive-class and negative-class mean vectors
   (reference: include/IOStructures.h:114-122). */
typedef struct svm_dataset{
    int l;
    int sparse;
    int maxdim;
    double *y;
    struct svm_sample **x;
    double *quadratic_value;
    struct svm_sample* features;
}svm_dataset;

/* reference: include/IOStructures.h:131 */
void freeDataset (svm_dataset data);
/* reference: include/IOStructures.h:140 */
void freeModel (model modelo);
/* reference: include/IOStructures.h:155 */
svm_dataset readTrainFile(char filename[]);
/* reference: include/IOStructures.h:171 */
svm_dataset readTrainFileCSV(char filename[],char* separator);
/* reference: include/IOStructures.h:187 */
svm_dataset readUnlabeledFile(char filename[]);
/* reference: include/IOStructures.h:205 */
svm_dataset readUnlabeledFileCSV(char filename[],char* separator);
/* reference: include/IOStructures.h:216 -- byte layout kept identical */
void storeModel(model * mod, FILE *Output);
/* reference: include/IOStructures.h:226 */
void readModel(model * mod, FILE *Input);
/* reference: include/IOStructures.h:238 */
void writeOutput (char fileoutput[], double *predictions, int size);

#ifdef __cplusplus
}
#endif

#endif
