/*
 * Scalar kernel functions (reference: src/kernels.c:63-183), exact FP64 CPU
 * versions for callers that evaluate single pairs.  The trainers and the
 * predictor never call these: they evaluate kernel tiles on the GPU.
 */
#include <math.h>
#include "kernels.h"

static double merge_dot(const svm_sample *a, const svm_sample *b) {
    double s = 0.0;
    while (a->index != -1 && b->index != -1) {
        if (a->index == b->index) {
            s += a->value * b->value;
            a++;
            b++;
        } else if (a->index > b->index) {
            b++;
        } else {
            a++;
        }
    }
    return s;
}

static double norm_form(const svm_sample *a, const svm_sample *b, double qa, double qb, double gamma) {
    double s = qa + qb;
    while (a->index != -1 && b->index != -1) {
        if (a->index == b->index) {
            s += -2.0 * a->value * b->value;
            a++;
            b++;
        } else if (a->index < b->index) {
            a++;
        } else {
            b++;
        }
    }
    return exp(-gamma * s);
}

double kernelFunction(svm_dataset dataset, int index1, int index2, properties props) {
    const svm_sample *a = dataset.x[index1], *b = dataset.x[index2];
    if (props.kernelType == 0) return merge_dot(a, b);
    if (dataset.sparse == 0) {
        double s = 0.0;
        for (int k = 0; k < dataset.maxdim; k++) {
            const double t = a[k].value - b[k].value;
            s += t * t;
        }
        return exp(-props.Kgamma * s);
    }
    return norm_form(a, b, dataset.quadratic_value[index1], dataset.quadratic_value[index2], props.Kgamma);
}

double kernelTest(svm_dataset dataset, int index1, model mymodel, int index2) {
    const svm_sample *a = dataset.x[index1], *b = mymodel.x[index2];
    if (mymodel.kernelType == 0) return merge_dot(a, b);
    return norm_form(a, b, dataset.quadratic_value[index1], mymodel.quadratic_value[index2], mymodel.Kgamma);
}
