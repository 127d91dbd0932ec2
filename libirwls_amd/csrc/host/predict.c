/*
 * Prediction (reference: src/LIBIRWLS-predict.c:56-224): decision values
 * bias + sum_j w_j k(x_i, sv_j) on the GPU by the fused kernel-block x vector
 * launch, with the stored norms of both sides as the reference's kernelTest
 * uses them (src/kernels.c:162-181).
 */
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "LIBIRWLS-predict.h"

/* The support vectors of a model viewed as a dataset. */
static svm_dataset sv_view(const model *m) {
    svm_dataset sv;
    memset(&sv, 0, sizeof(sv));
    sv.l = m->nSVs;
    sv.sparse = m->sparse;
    sv.maxdim = m->maxdim;
    sv.x = m->x;
    sv.quadratic_value = m->quadratic_value;
    sv.features = m->features;
    return sv;
}

/* Feature range of the rows to predict: that of their pinned device copy,
   or a scan of the rows. */
static void test_range(const svm_dataset *ds, int *base, int *maxindex) {
    if (irwls_devdata_pinned_range(ds, base, maxindex)) return;
    *base = irwls_dataset_base(ds, ds->l);
    *maxindex = irwls_dataset_maxindex(ds, ds->l);
}

static double *decision_values(svm_dataset dataset, model m) {
    const int n = dataset.l;
    double *out = (double *)irwls_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    if (n <= 0) return out;
    if (m.nSVs <= 0) {
        for (int i = 0; i < n; i++) out[i] = m.bias;
        return out;
    }
    irwls_dev_init();
    svm_dataset sv = sv_view(&m);
    int bt, mt;
    test_range(&dataset, &bt, &mt);
    const int bm = irwls_dataset_base(&sv, m.nSVs);
    const int base = (bt < bm) ? bt : bm;
    const int mm = irwls_dataset_maxindex(&sv, m.nSVs);
    const int cols = ((mt > mm) ? mt : mm) + 1 - base;
    irwls_devdata *dt = irwls_devdata_acquire(&dataset, 0, base, cols);
    irwls_devdata *dm = irwls_devdata_acquire(&sv, 0, base, cols);
    const int d = (dt->d > dm->d) ? dt->d : dm->d;
    irwls_rows r = irwls_devdata_rows(dt, NULL, n);
    irwls_rows c = irwls_devdata_rows(dm, NULL, m.nSVs);
    r.nrm = dt->qv;
    c.nrm = dm->qv;
    if (r.X) r.ldx = dt->d;
    if (c.X) c.ldx = dm->d;
    double *d_w = (double *)irwls_dev_alloc((size_t)m.nSVs * sizeof(double));
    double *d_out = (double *)irwls_dev_alloc((size_t)n * sizeof(double));
    irwls_dev_h2d(d_w, m.weights, (size_t)m.nSVs * sizeof(double));
    irwls_dev_kernel_matvec(&r, &c, d, m.kernelType, m.Kgamma, d_w, 0.0, 1.0, m.bias, NULL, d_out);
    irwls_dev_d2h(out, d_out, (size_t)n * sizeof(double));
    irwls_dev_free(d_w);
    irwls_dev_free(d_out);
    irwls_devdata_release(dm);
    irwls_devdata_release(dt);
    return out;
}

static void print_accuracy(const svm_dataset *ds, const double *pred) {
    double hits = 0.0;
    for (int i = 0; i < ds->l; i++) {
        if (pred[i] > 0 && ds->y[i] > 0) hits++;
        if (pred[i] <= 0 && ds->y[i] <= 0) hits++;
    }
    printf("Accuracy: %f\n", hits / (double)ds->l);
}

int softTestMany(svm_dataset dataset, const model *models, int nmodels, predictProperties props, double *out) {
    const int n = dataset.l;
    if (n <= 0 || nmodels <= 0) return 0;
    irwls_dev_init();
    int bt, mt;
    test_range(&dataset, &bt, &mt);
    double *d_out = (double *)irwls_dev_alloc((size_t)n * nmodels * sizeof(double));
    /* kept until the single read-back, so no release waits on a queued launch */
    irwls_devdata **dt = (irwls_devdata **)irwls_xcalloc((size_t)nmodels, sizeof(irwls_devdata *));
    irwls_devdata **dm = (irwls_devdata **)irwls_xcalloc((size_t)nmodels, sizeof(irwls_devdata *));
    double **d_w = (double **)irwls_xcalloc((size_t)nmodels, sizeof(double *));
    for (int g = 0; g < nmodels; g++) {
        const model *m = &models[g];
        double *og = d_out + (size_t)g * n;
        if (m->nSVs <= 0) {
            irwls_rows r, none;
            memset(&r, 0, sizeof(r));
            memset(&none, 0, sizeof(none));
            r.n = n;
            irwls_dev_kernel_matvec(&r, &none, 0, m->kernelType, m->Kgamma, NULL, 0.0, 1.0, m->bias, NULL, og);
            continue;
        }
        /* the same column layout softTest uses for this model (the device
           copy of the test rows is shared between models that agree on it) */
        svm_dataset sv = sv_view(m);
        const int bm = irwls_dataset_base(&sv, sv.l), mm = irwls_dataset_maxindex(&sv, sv.l);
        const int base = (bt < bm) ? bt : bm;
        const int cols = ((mt > mm) ? mt : mm) + 1 - base;
        dt[g] = irwls_devdata_acquire(&dataset, 0, base, cols);
        dm[g] = irwls_devdata_acquire(&sv, 0, base, cols);
        const int d = (dt[g]->d > dm[g]->d) ? dt[g]->d : dm[g]->d;
        irwls_rows r = irwls_devdata_rows(dt[g], NULL, n);
        irwls_rows c = irwls_devdata_rows(dm[g], NULL, m->nSVs);
        r.nrm = dt[g]->qv;
        c.nrm = dm[g]->qv;
        if (r.X) r.ldx = dt[g]->d;
        if (c.X) c.ldx = dm[g]->d;
        d_w[g] = (double *)irwls_dev_alloc((size_t)m->nSVs * sizeof(double));
        irwls_dev_h2d(d_w[g], m->weights, (size_t)m->nSVs * sizeof(double));
        irwls_dev_kernel_matvec(&r, &c, d, m->kernelType, m->Kgamma, d_w[g], 0.0, 1.0, m->bias, NULL, og);
    }
    irwls_dev_d2h(out, d_out, (size_t)n * nmodels * sizeof(double));
    for (int g = 0; g < nmodels; g++) {
        if (d_w[g]) irwls_dev_free(d_w[g]);
        if (dm[g]) irwls_devdata_release(dm[g]);
        if (dt[g]) irwls_devdata_release(dt[g]);
    }
    free(d_w);
    free(dm);
    free(dt);
    irwls_dev_free(d_out);
    if (props.Labels == 1)
        for (int g = 0; g < nmodels; g++) print_accuracy(&dataset, out + (size_t)g * n);
    return 0;
}

double *softTest(svm_dataset dataset, model mymodel, predictProperties props) {
    double *p = decision_values(dataset, mymodel);
    if (props.Labels == 1) print_accuracy(&dataset, p);
    return p;
}

double *test(svm_dataset dataset, model mymodel, predictProperties props) {
    double *p = decision_values(dataset, mymodel);
    for (int i = 0; i < dataset.l; i++) p[i] = (p[i] >= 0.0) ? 1.0 : -1.0;
    if (props.Labels == 1) print_accuracy(&dataset, p);
    return p;
}

void printPredictInstructions(void) {
    fprintf(stderr, "LIBIRWLS-predict: This software predicts the label of a SVM given a data set of samples and a model obtained with full-train or budgeted-train");
    fprintf(stderr, "and store the results in an output file.\n\n");
    fprintf(stderr, "Usage: LIBIRWLS-predict [options] data_set_file model_file output_file\n\n");
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "  -t Number of Threads: (default 1)\n");
    fprintf(stderr, "  -l type of data set: (default 0)\n");
    fprintf(stderr, "       0 -- Data set with no target as first dimension.\n");
    fprintf(stderr, "       1 -- Data set with label as first dimension (obtains accuracy too)\n");
    fprintf(stderr, "  -s Soft output: (default 0)\n");
    fprintf(stderr, "       0 -- Obtains the class of every data (It takes values of +1 or -1).\n");
    fprintf(stderr, "       1 -- The output before the class decision (Useful to combine in ensembles with other algorithms).\n");
    fprintf(stderr, "  -f file format: (default 1)\n");
    fprintf(stderr, "       0 -- CSV format (comma separator)\n");
    fprintf(stderr, "       1 -- libsvm format\n");
    fprintf(stderr, "  -p separator: csv separator character (default \",\" if csv format is selected)\n");
    fprintf(stderr, "  -v verbose: (default 1)\n");
    fprintf(stderr, "       0 -- No screen messages\n");
    fprintf(stderr, "       1 -- Screen messages\n");
}

predictProperties parsePredictParameters(int *argc, char ***argv) {
    predictProperties props;
    memset(&props, 0, sizeof(props));
    props.Labels = 0;
    props.Threads = 1;
    props.Soft = 0;
    props.file = 1;
    props.separator = ",";
    props.verbose = 1;
    int i;
    for (i = 1; i < *argc; ++i) {
        if ((*argv)[i][0] != '-') break;
        if (++i >= *argc) {
            printPredictInstructions();
            exit(1);
        }
        const char *name = &(*argv)[i - 1][1];
        char *val = (*argv)[i];
        if (strcmp(name, "t") == 0) props.Threads = (int)atof(val);
        else if (strcmp(name, "f") == 0) props.file = atoi(val);
        else if (strcmp(name, "p") == 0) props.separator = val;
        else if (strcmp(name, "v") == 0) props.verbose = atoi(val);
        else if (strcmp(name, "l") == 0) {
            props.Labels = atoi(val);
            if (props.Labels != 0 && props.Labels != 1) {
                printf("\nInvalid type of test data set:%d\n", props.Labels);
                exit(2);
            }
        } else if (strcmp(name, "s") == 0) {
            props.Soft = atoi(val);
            if (props.Soft != 0 && props.Soft != 1) {
                printf("\nInvalid type of output (-s param):%d\n", props.Soft);
                exit(2);
            }
        } else {
            fprintf(stderr, "Unknown parameter %s\n", name);
            printPredictInstructions();
            exit(2);
        }
    }
    for (int j = 1; i + j - 1 < *argc; ++j) (*argv)[j] = (*argv)[i + j - 1];
    *argc -= i - 1;
    return props;
}
