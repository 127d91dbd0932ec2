/*
 * Internal declarations shared by the C host files of libirwls.
 */
#ifndef IRWLS_HOST_INTERNAL_H
#define IRWLS_HOST_INTERNAL_H

#include <stdint.h>
#include <stdlib.h>
#include "IOStructures.h"
#include "irwls_device.h"
#include "irwls_gpu.h"

/* A dataset resident in device memory.  Feature index f maps to column
   f - base (base 1 when every index is >= 1, the libsvm convention, else 0).
   Dense (svm_dataset.sparse == 0): X is rows x d row-major.  For well-formed
   dense rows (every row lists the same indices 1..maxdim) this equals the
   reference's positional walk (src/kernels.c:90-97); nrm = |x|^2 computed on
   the device in feature order.
   CSR (sparse == 1): rowptr/col/val; nrm = the dataset's stored
   quadratic_value, which the reference's sparse kernel uses
   (src/kernels.c:101), including the class-mean rows' stored norms.
   qv always holds the stored quadratic_value (prediction uses it,
   src/kernels.c:163). */
typedef struct irwls_devdata {
    /* identity: the dataset's features, x, y and quadratic_value arrays */
    const svm_sample *key;
    const void *key_x, *key_y, *key_q;
    /* pinned copies: fingerprints of the host's y and quadratic_value
       (rows [0, rows)) and of the mean rows' entries, checked on every
       acquire so that labels or mean rows changed in place are uploaded
       again */
    uint64_t fp_yq, fp_mean;
    int l;                     /* labelled rows */
    int rows;                  /* uploaded rows (l, or l+2 with the mean rows) */
    int d;                     /* column count of the device layout */
    int dense;
    int base;                  /* feature index of column 0 */
    double *X;
    int64_t *rowptr;
    int32_t *col;
    double *val;
    double *nrm;
    double *qv;
    double *y;
    int pinned;                /* kept resident between calls */
    double nnz_per_row;        /* CSR: mean nonzeros per uploaded row */
    /* mixed precision (dense only, built on first use): BF16 hi/lo rows,
       ldb = d rounded up to 32, FP32 norms of the split representation */
    uint16_t *Xh, *Xl;
    int ldb;
    float *nrm32;
} irwls_devdata;

/* 1 when every feature index of rows [0, rows) is >= 1, else 0. */
int irwls_dataset_base(const svm_dataset *ds, int rows);
/* Largest feature index of rows [0, rows). */
int irwls_dataset_maxindex(const svm_dataset *ds, int rows);

/* Upload (or find the pinned copy of) a dataset.  with_means asks for the two
   class-mean rows l and l+1 as well; base as above (-1 = derive it);
   min_cols widens the column space (prediction aligns data and model). */
/* Host parallel loops (par.c): fn(arg, t, begin, end) over nt contiguous
   ranges of [0, n), the caller's thread taking t = 0; irwls_host_threads
   gives min(16, cores, work / grain), at least 1. */
typedef void (*irwls_range_fn)(void *arg, int t, long begin, long end);
void irwls_parallel_ranges(long n, int nt, irwls_range_fn fn, void *arg);
int irwls_host_threads(long work, long grain);
/* 1 on a worker thread of trainFULLBatched */
int irwls_batch_worker(void);

irwls_devdata *irwls_devdata_acquire(const svm_dataset *ds, int with_means, int base, int min_cols);
/* Base and largest feature index of a pinned dataset's device copy (its
   column range, which may extend past the largest index present: those
   columns are zero); 0 when the dataset is not pinned.  Spares a caller the
   scan of every row that irwls_dataset_base and irwls_dataset_maxindex make. */
int irwls_devdata_pinned_range(const svm_dataset *ds, int *base, int *maxindex);
void irwls_devdata_release(irwls_devdata *dd);
/* Row view of the whole device dataset (idx may select rows). */
irwls_rows irwls_devdata_rows(const irwls_devdata *dd, const int32_t *d_idx, int n);

/* Kernel block x vector and materialised kernel block over rows of a device
   dataset (row selections d_ridx / d_cidx may be NULL), in the current
   precision: BF16 hi/lo products for dense data in mixed mode, FP64
   otherwise.  Arguments as irwls_dev_kernel_matvec / _matrix. */
void irwls_kmv(irwls_devdata *dd, const int32_t *d_ridx, int rn, const int32_t *d_cidx, int cn, int kernel,
               double gamma, const double *d_w, double a, double b, double cst, const double *d_rs, double *d_out);
void irwls_kmat(irwls_devdata *dd, const int32_t *d_ridx, int rn, const int32_t *d_cidx, int cn, int kernel,
                double gamma, double *d_K, int ldk);

/* Current options (environment defaults merged with irwls_gpu_set_options). */
const irwls_gpu_options *irwls_opts(void);

/* Host wall clock in seconds. */
double irwls_now(void);

/* Device work buffers of the full trainer's working-set IRWLS. */
typedef struct irwls_ws {
    int cap;                   /* capacity in working-set rows */
    double *Kw, *H, *esub, *G13;
    /* The small per-iteration vectors live in three contiguous device spans,
       each uploaded with one copy from a host mirror of the same layout
       (irwls_ws_upload): span 0, the KKT system (S1, ys, as, rhs); span 1,
       the error update (cols, w); span 2, G13 (S1g, S3, wg, rs).  w also
       receives the dense solve's solution. */
    double *pk, *hpk;
    size_t span_off[3], span_len[3];      /* in doubles */
    int32_t *S1, *cols, *S1g, *S3;
    double *ys, *as, *rhs, *w, *wg, *rs;
    int32_t *hS1, *hcols, *hS1g, *hS3;
    double *hys, *has, *hrhs, *hw, *hwg, *hrs;
    /* device state of the inner loop: the solve's solution (x), the labels
       (yl), GIN, the weights (beta) and the weights after a solve (bnew) of
       the working set; nc_cst holds the changed-column count (int32 at
       word 0) and the error update's constant (word 1) */
    double *x, *yl, *gin, *beta, *bnew, *nc_cst;
} irwls_ws;

void irwls_ws_reserve(irwls_ws *ws, int cap);
void irwls_ws_free(irwls_ws *ws);
void irwls_ws_upload(irwls_ws *ws, int span);

/* Inner IRWLS on a working set whose rows are `wrows` (device view).  Same
   contract as subIRWLS (src/full-train.c:96-362): y = labels of the working
   set (host), GIN has l+1 entries, e and beta (l and l+1 entries) are updated
   in place; returns betaBest (malloc, l+1). */
double *irwls_sub_irwls_dev(irwls_ws *ws, irwls_devdata *dd, const int32_t *d_widx, int l, const double *y,
                            properties props, const double *GIN, double *e, double *beta);

/* Mixed precision (mixed_kkt.c).  A new generation marks a new kernel
   block (one per subIRWLS call): factors of an older one are never reused. */
long irwls_mixed_new_generation(void);
int irwls_kkt_solve_mixed(const double *d_H, int nS1, const int32_t *S1, long gen, const double *d_rhs,
                          const double *d_ys, double last, double *d_x, irwls_gpu_stats *st);
void irwls_mixed_release(void);

/* Learned SGMA proposal (proposal.c). */
typedef struct irwls_proposal irwls_proposal;
irwls_proposal *irwls_proposal_begin(int n, int B, int pmax, int log_cap);
void irwls_proposal_end(irwls_proposal *lp);
int irwls_proposal_trained(const irwls_proposal *lp);
void irwls_proposal_add_centroid(irwls_proposal *lp, int s, const double *d_col);
void irwls_proposal_add_centroid_rows(irwls_proposal *lp, irwls_devdata *dd, properties props, int s,
                                      const int32_t *d_c);
void irwls_proposal_features(irwls_proposal *lp, irwls_devdata *dd, const svm_dataset *ds, properties props,
                             int s, int P, const int32_t *d_c, const int32_t *cidx, const double *d_KNC,
                             const double *d_Z, const double *d_eta, float pos_frac, const int32_t *d_npos);
void irwls_proposal_log(irwls_proposal *lp, int P, const double *q, const double *eta);
int irwls_proposal_train(irwls_proposal *lp, int steps, float lr, irwls_gpu_stats *st);
void irwls_proposal_score(irwls_proposal *lp, int P, float *score);
const float *irwls_proposal_score_dev(irwls_proposal *lp, int P);

/* Batched training (batch.c).  irwls_rand is libc rand(), or the calling
   worker's private glibc-compatible stream in a batched run.
   irwls_batch_inner joins a batched run's pending inner iterations (all
   operands on the device, irwls_inner_work) and returns 1 when done, or
   returns 0 outside a batched run. */
int irwls_rand(void);
/* A run of libc rand() draws without libc's per-call lock: the same values
   and the same libc state afterwards as calling rand() (irwls_rand() on a
   batched worker) once per draw.  No other thread may call rand() between
   begin and end. */
typedef struct {
    _Alignas(16) char opaque[448];   /* the generator state and two state buffers (batch.c) */
} irwls_rand_block;
void irwls_rand_begin(irwls_rand_block *b);
int irwls_rand_draw(irwls_rand_block *b);
void irwls_rand_end(irwls_rand_block *b);
int irwls_batch_inner(const irwls_inner_work *w);
/* Marks the end of the calling worker's inner loop (no-op outside a batched
   run): batches stop waiting for it. */
void irwls_batch_inner_end(void);

/* Counters reported by irwls_gpu_last_stats (the last training call). */
void irwls_stats_set(const irwls_gpu_stats *s);

/* libc-compatible allocation with exit(2) on failure. */
void *irwls_xmalloc(size_t bytes);
void *irwls_xcalloc(size_t n, size_t sz);

#endif
