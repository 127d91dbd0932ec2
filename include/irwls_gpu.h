/*
 * Additive C API of libirwls (no counterpart in the reference).
 *
 * The reference entry points keep their prototypes and struct layouts; the
 * knobs of the GPU build live here instead, so sizeof(properties) does not
 * change.  Every function is optional: a caller that only uses the reference
 * interface gets the plain FP64 path with its defaults.
 */
#ifndef IRWLS_GPU_H
#define IRWLS_GPU_H

#include "IOStructures.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct irwls_gpu_options {
    int precision;        /* 0 = FP64 (default), 1 = mixed (BF16x3 kernel GEMM, FP32 Cholesky + refinement);
                             2 = only the mixed solver, 3 = only the BF16x3 kernel GEMM (ablations) */
    int sgma_candidates;  /* SGMA candidates per step (default 64, as the reference) */
    int proposal;         /* 0 = plain SGMA draws (default), 1 = learned shortlist */
    int proposal_keep;    /* learned proposal: candidates kept for exact scoring (0: the network's choice) */
    double refine_tol;    /* mixed precision: relative residual target (default 1e-10) */
    int proposal_pool;    /* learned proposal: candidates drawn per step (0: sgma_candidates, the draws of plain
                             SGMA, which the network shortlists; at most 1024) */
    int proposal_warmup;  /* learned proposal: plain SGMA steps that train the network (0: B/10 in [20, 100]) */
    int proposal_rows;    /* learned proposal: rows on which the shortlist's gains are computed (0: all n, the
                             exact gains; m < n: a fixed random subset of m rows, an estimate) */
    int par_max_iters;    /* IRWLSpar: stop after this many iterations (0: the reference's rules only; a
                             diagnostic cap to compare iterates) */
} irwls_gpu_options;

/* Defaults, overridable by IRWLS_PRECISION=fp64|mixed, IRWLS_SGMA_CAND=<n>,
   IRWLS_PROPOSAL=sgma|learned. */
void irwls_gpu_get_options(irwls_gpu_options *out);
void irwls_gpu_set_options(const irwls_gpu_options *opts);

/* Keep a dataset resident in device memory so later trainFULL / SGMA /
   IRWLSpar / test calls on it skip the host-to-device upload.  The dataset is
   identified by its features, x, y and quadratic_value pointers.  Labels,
   norms and the mean rows may be changed in place (one-vs-rest relabelling):
   every call compares their fingerprints and uploads what changed.  Feature
   values of the other rows must not change while pinned.  freeDataset unpins. */
void irwls_gpu_pin_dataset(svm_dataset dataset, int with_means);
void irwls_gpu_unpin_dataset(svm_dataset dataset);

/* Per-kernel timers (HIP events on the library stream). */
void irwls_gpu_timing(int enable);
/* Restrict timing to the kinds whose bit is set (bit k = timer kind k; default all). */
void irwls_gpu_timing_mask(unsigned mask);
/* Time only every stride-th launch of each timed kind (1, the default:
   every launch).  The events add a few microseconds to each timed launch;
   a stride keeps that out of the measured run while the sampled launches
   still give the kind's average launch duration. */
void irwls_gpu_timing_stride(int stride);
void irwls_gpu_timer(int kind, long *launches, double *ms, double *flops, double *bytes);

/* Counters of the last training call. */
typedef struct irwls_gpu_stats {
    int outer_iters;
    int inner_iters;
    int solves;
    int sgma_steps;
    int irwlspar_iters;
    double t_total;        /* seconds, host wall clock */
    double t_solve;        /* seconds inside the dense solver */
    int factorizations;    /* Cholesky factorizations (mixed precision: FP32 factors) */
    int pcg_iters;         /* mixed precision: CG iterations over all solves */
    int fp64_fallbacks;    /* mixed precision: solves done by the FP64 path instead */
    int learned_steps;     /* SGMA steps chosen through the learned proposal */
    double mlp_loss0;      /* proposal network: loss before / after its fit */
    double mlp_loss;
    int batch_rounds;      /* trainFULLBatched: batched solve rounds */
    double t_wait;         /* seconds the host waited for the device (read-backs, synchronisations;
                              trainFULLBatched: summed over the workers, batched solves excluded) */
    int par_m_decays;      /* IRWLSpar: times its weight cap M was divided by 10 */
} irwls_gpu_stats;
void irwls_gpu_last_stats(irwls_gpu_stats *out);

/* In-memory dataset builders with the reference readers' semantics
   (src/IOStructures.c:88-275): the two class-mean rows are appended and
   |x|^2 is stored.  Dense rows list every index 1..d; CSR columns are
   0-based (stored as index col+1).  Free with freeDataset. */
svm_dataset irwls_dataset_from_dense(const double *X, const double *y, int n, int d);
svm_dataset irwls_dataset_from_csr(const long long *rowptr, const int *col, const double *val,
                                   const double *y, int n);

/* Dataset from a dense row-major n x d matrix with the conversion of the
   reference's Python extension (numpy2dataset, and numpy2datasetWithAverage
   when with_average is nonzero; python-package/pythonmodule.c:210-330): only
   nonzero values stored, sparse = 1 as soon as one value is nonzero, and,
   with with_average, the extension's two mean rows exactly as it builds them
   (the first lists indices 1..d and an index-0 entry, the second is empty,
   their norms are sums of the mean values).  Backs the Python 3 module
   libirwls_amd.LIBIRWLS.  Free with freeDataset. */
svm_dataset irwls_dataset_from_numpy(const double *X, const double *y, int n, int d, int with_average);

/* Binary dataset cache (no counterpart in the reference; a parsed text
   dataset saved once reloads without parsing).  irwls_dataset_save writes
   the dataset's rows (with the two class-mean rows when with_means is
   nonzero, as the labelled readers build them) and returns 0, or -1 when
   the file cannot be written.  irwls_dataset_load returns the dataset as
   it was saved; a file that is not such a cache is reported as "Wrong file
   format" with exit(2), as the text readers do.  Free with freeDataset. */
int irwls_dataset_save(const svm_dataset *ds, int with_means, const char *path);
svm_dataset irwls_dataset_load(const char *path);

/* Batched full training: nprob independent problems (datasets[p],
   props[p]); out[p] receives trainFULL's result for problem p (l_p + 1
   doubles, malloc'd, the caller frees it).  Each problem equals its
   standalone trainFULL run after srand(0) (the reference's full-train main),
   bit for bit; their dense solves run batched.  Returns 0.  No counterpart
   in the reference. */
int trainFULLBatched(const svm_dataset *datasets, const properties *props, int nprob, double **out);

/* Decision values of nmodels models on one test set (no counterpart in the
   reference; one-vs-rest bundles and grids of models): out[g * l + i] is
   what softTest(dataset, models[g], ...) returns for row i, bit for bit.
   The test rows are uploaded once and the models' launches are queued
   without waiting in between; out holds l * nmodels doubles (caller's
   memory).  With props.Labels == 1 each model's accuracy is printed as
   softTest prints it.  Returns 0. */
int softTestMany(svm_dataset dataset, const model *models, int nmodels, predictProperties props, double *out);

/* Release the device context (initMemory/freeMemory do this for callers of
   the reference interface). */
void irwls_gpu_shutdown(void);

#ifdef __cplusplus
}
#endif

#endif
