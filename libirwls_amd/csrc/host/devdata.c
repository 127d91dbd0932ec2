/*
 * Device residency of datasets, options, timers and small host utilities.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>
#include <time.h>
#include "internal.h"

double irwls_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void *irwls_xmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "libirwls: out of host memory (%zu bytes)\n", bytes);
        exit(2);
    }
    return p;
}

void *irwls_xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) {
        fprintf(stderr, "libirwls: out of host memory (%zu x %zu bytes)\n", n, sz);
        exit(2);
    }
    return p;
}

/* ---- options ----------------------------------------------------------- */
static irwls_gpu_options g_opts;
static int g_opts_init = 0;

static void opts_defaults(void) {
    if (g_opts_init) return;
    g_opts.precision = 0;
    g_opts.sgma_candidates = 64;
    g_opts.proposal = 0;
    g_opts.proposal_keep = 8;
    g_opts.refine_tol = 1e-10;
    g_opts.proposal_pool = 0;
    g_opts.proposal_warmup = 0;
    g_opts.proposal_rows = 0;
    g_opts.par_max_iters = 0;
    const char *p = getenv("IRWLS_PRECISION");
    if (p && strcasecmp(p, "mixed") == 0) g_opts.precision = 1;
    const char *c = getenv("IRWLS_SGMA_CAND");
    if (c && atoi(c) > 0) g_opts.sgma_candidates = atoi(c);
    const char *q = getenv("IRWLS_PROPOSAL");
    if (q && strcasecmp(q, "learned") == 0) g_opts.proposal = 1;
    g_opts_init = 1;
}

const irwls_gpu_options *irwls_opts(void) {
    opts_defaults();
    return &g_opts;
}

void irwls_gpu_get_options(irwls_gpu_options *out) {
    opts_defaults();
    *out = g_opts;
}

void irwls_gpu_set_options(const irwls_gpu_options *o) {
    opts_defaults();
    g_opts = *o;
    if (g_opts.sgma_candidates < 1) g_opts.sgma_candidates = 1;
    if (g_opts.sgma_candidates > 1024) g_opts.sgma_candidates = 1024;
    if (g_opts.proposal_keep < 0) g_opts.proposal_keep = 0;
    if (g_opts.proposal_pool < 0) g_opts.proposal_pool = 0;
    if (g_opts.proposal_pool > 1024) g_opts.proposal_pool = 1024;
    if (g_opts.proposal_warmup < 0) g_opts.proposal_warmup = 0;
    if (g_opts.par_max_iters < 0) g_opts.par_max_iters = 0;
}

void irwls_gpu_timing(int enable) { irwls_dev_timing(enable); }
void irwls_gpu_timing_mask(unsigned mask) { irwls_dev_timing_mask(mask); }
void irwls_gpu_timing_stride(int stride) { irwls_dev_timing_stride(stride); }

void irwls_gpu_timer(int kind, long *launches, double *ms, double *flops, double *bytes) {
    irwls_dev_timer_get(kind, launches, ms, flops, bytes);
}

/* ---- residency --------------------------------------------------------- */
#define MAX_PINNED 16
static irwls_devdata *g_pinned[MAX_PINNED];
/* The dataset table and the lazily built device copies are shared by the
   worker threads of a batched run: changes happen under this lock, and a
   thread finishes (synchronises) the device work of a build before it lets
   the others see the result. */
static pthread_mutex_t g_dd_mu = PTHREAD_MUTEX_INITIALIZER;

int irwls_dataset_base(const svm_dataset *ds, int rows) {
    for (int i = 0; i < rows; i++)
        for (const svm_sample *s = ds->x[i]; s->index != -1; s++)
            if (s->index < 1) return 0;
    return 1;
}

int irwls_dataset_maxindex(const svm_dataset *ds, int rows) {
    int m = 0;
    for (int i = 0; i < rows; i++)
        for (const svm_sample *s = ds->x[i]; s->index != -1; s++)
            if (s->index > m) m = s->index;
    return m;
}

/* 64-bit FNV-1a over 8-byte words. */
static uint64_t fp_words(uint64_t h, const void *p, size_t nwords) {
    const uint64_t *w = (const uint64_t *)p;
    for (size_t i = 0; i < nwords; i++) {
        h ^= w[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

static uint64_t fp_labels(const svm_dataset *ds, int rows) {
    uint64_t h = 0xcbf29ce484222325ULL;
    if (ds->y) h = fp_words(h, ds->y, (size_t)rows);
    if (ds->quadratic_value) h = fp_words(h, ds->quadratic_value, (size_t)rows);
    return h;
}

static uint64_t fp_means(const svm_dataset *ds, int rows) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (int i = ds->l; i < rows; i++) {
        const svm_sample *s = ds->x[i];
        size_t k = 0;
        while (s[k].index != -1) k++;
        h = fp_words(h, &k, 1);
        for (size_t j = 0; j < k; j++) {
            const uint64_t idx = (uint64_t)(uint32_t)s[j].index;
            h = fp_words(h, &idx, 1);
            h = fp_words(h, &s[j].value, 1);
        }
    }
    return h;
}

static void upload_fail(int row, int index) {
    fprintf(stderr, "libirwls: feature index %d of row %d is out of range\n", index, row);
    exit(2);
}

/* ---- upload: the AoS rows walked on several host threads, written into
   pinned chunks whose transfers overlap the filling of the next chunk ---- */

typedef struct {
    const svm_dataset *ds;
    int below1[16], maxidx[16];
} scan_job;

/* irwls_dataset_base and irwls_dataset_maxindex in one pass */
static void scan_rows(void *arg, int t, long b, long e) {
    scan_job *j = (scan_job *)arg;
    int below = 0, m = 0;
    for (long i = b; i < e; i++)
        for (const svm_sample *s = j->ds->x[i]; s->index != -1; s++) {
            if (s->index < 1) below = 1;
            if (s->index > m) m = s->index;
        }
    j->below1[t] = below;
    j->maxidx[t] = m;
}

typedef struct {
    const svm_dataset *ds;
    int base, d;
    size_t row_bytes;
    double *buf;      /* the chunk being filled */
    long r0;          /* its first row */
} dense_job;

static void fill_dense_rows(void *arg, int t, long b, long e) {
    (void)t;
    dense_job *j = (dense_job *)arg;
    for (long i = b; i < e; i++) {
        double *row = j->buf + (size_t)(i - j->r0) * j->d;
        memset(row, 0, j->row_bytes);
        for (const svm_sample *s = j->ds->x[i]; s->index != -1; s++) {
            const int c = s->index - j->base;
            if (c < 0) upload_fail((int)i, s->index);
            if (c < j->d) row[c] = s->value;
        }
    }
}

static void fill_dense_rows_rel(void *arg, int t, long b, long e) {
    dense_job *j = (dense_job *)arg;
    fill_dense_rows(arg, t, j->r0 + b, j->r0 + e);
}

static void fill_dense(void *arg, void *buf, size_t off, size_t len) {
    dense_job *j = (dense_job *)arg;
    j->buf = (double *)buf;
    j->r0 = (long)(off / j->row_bytes);
    const long nr = (long)(len / j->row_bytes);
    irwls_parallel_ranges(nr, irwls_host_threads(nr * (long)j->d, 1L << 17), fill_dense_rows_rel, j);
}

typedef struct {
    const svm_dataset *ds;
    int base;
    const int64_t *rp;
    int rows;
    int64_t *len;      /* per row: entry count (count pass) */
    void *buf;         /* the chunk being filled */
    int64_t k0, k1;    /* its entries */
    int what;          /* 0: columns (int32), 1: values (double) */
} csr_job;

static void count_rows(void *arg, int t, long b, long e) {
    (void)t;
    csr_job *j = (csr_job *)arg;
    for (long i = b; i < e; i++) {
        int64_t c = 0;
        for (const svm_sample *s = j->ds->x[i]; s->index != -1; s++) {
            if (s->index < j->base) upload_fail((int)i, s->index);
            c++;
        }
        j->len[i] = c;
    }
}

/* the row holding entry k: the last row whose start is <= k */
static int row_of(const int64_t *rp, int rows, int64_t k) {
    int lo = 0, hi = rows - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rp[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

static void fill_csr_part(void *arg, int t, long b, long e) {
    (void)t;
    csr_job *j = (csr_job *)arg;
    const int64_t k0 = j->k0 + b, k1 = j->k0 + e;
    if (k0 >= k1) return;
    int i = row_of(j->rp, j->rows, k0);
    int64_t k = k0;
    while (k < k1) {
        const svm_sample *s = j->ds->x[i] + (k - j->rp[i]);
        const int64_t stop = j->rp[i + 1] < k1 ? j->rp[i + 1] : k1;
        for (; k < stop; k++, s++) {
            if (j->what == 0) ((int32_t *)j->buf)[k - j->k0] = s->index - j->base;
            else ((double *)j->buf)[k - j->k0] = s->value;
        }
        i++;
    }
}

static void fill_csr(void *arg, void *buf, size_t off, size_t len) {
    csr_job *j = (csr_job *)arg;
    const size_t es = j->what == 0 ? sizeof(int32_t) : sizeof(double);
    j->buf = buf;
    j->k0 = (int64_t)(off / es);
    j->k1 = j->k0 + (int64_t)(len / es);
    const long ne = (long)(j->k1 - j->k0);
    irwls_parallel_ranges(ne, irwls_host_threads(ne, 1L << 18), fill_csr_part, j);
}

static const size_t kUploadChunk = (size_t)32 << 20;

static irwls_devdata *upload(const svm_dataset *ds, int with_means, int base, int min_cols) {
    irwls_devdata *dd = (irwls_devdata *)irwls_xcalloc(1, sizeof(irwls_devdata));
    dd->key = ds->features;
    dd->key_x = ds->x;
    dd->key_y = ds->y;
    dd->key_q = ds->quadratic_value;
    dd->l = ds->l;
    dd->rows = ds->l + (with_means ? 2 : 0);
    dd->dense = (ds->sparse == 0);
    const int rows = dd->rows;
    scan_job sj;
    memset(&sj, 0, sizeof(sj));
    sj.ds = ds;
    const int nts = irwls_host_threads(rows, 4096);
    irwls_parallel_ranges(rows, nts, scan_rows, &sj);
    int below1 = 0, maxidx = 0;
    for (int t = 0; t < nts; t++) {
        below1 |= sj.below1[t];
        if (sj.maxidx[t] > maxidx) maxidx = sj.maxidx[t];
    }
    if (base < 0) base = below1 ? 0 : 1;
    dd->base = base;
    int d = maxidx + 1 - base;
    if (d < min_cols) d = min_cols;
    if (d < 1) d = 1;
    dd->d = d;
    const size_t rr = (size_t)(rows > 0 ? rows : 1);
    double *hy = (double *)irwls_xcalloc(rr, sizeof(double));
    for (int i = 0; i < rows; i++) hy[i] = ds->y ? ds->y[i] : 0.0;
    dd->y = (double *)irwls_dev_alloc(rr * sizeof(double));
    irwls_dev_h2d(dd->y, hy, (size_t)rows * sizeof(double));
    free(hy);
    dd->qv = (double *)irwls_dev_alloc(rr * sizeof(double));
    if (ds->quadratic_value) irwls_dev_h2d(dd->qv, ds->quadratic_value, (size_t)rows * sizeof(double));
    if (dd->dense) {
        dd->X = (double *)irwls_dev_alloc((size_t)rows * d * sizeof(double));
        dense_job j;
        memset(&j, 0, sizeof(j));
        j.ds = ds;
        j.base = base;
        j.d = d;
        j.row_bytes = (size_t)d * sizeof(double);
        size_t per = kUploadChunk / j.row_bytes;
        if (per < 1) per = 1;
        irwls_dev_h2d_chunked(dd->X, (size_t)rows * j.row_bytes, per * j.row_bytes, fill_dense, &j);
        dd->nrm = (double *)irwls_dev_alloc(rr * sizeof(double));
        irwls_dev_row_norms(dd->X, d, rows, d, dd->nrm);
    } else {
        int64_t *rp = (int64_t *)irwls_xmalloc((size_t)(rows + 1) * sizeof(int64_t));
        csr_job j;
        memset(&j, 0, sizeof(j));
        j.ds = ds;
        j.base = base;
        j.rows = rows;
        j.len = rp + 1;
        irwls_parallel_ranges(rows, irwls_host_threads(rows, 4096), count_rows, &j);
        rp[0] = 0;
        for (int i = 0; i < rows; i++) rp[i + 1] += rp[i];
        const int64_t nnz = rp[rows];
        j.rp = rp;
        dd->rowptr = (int64_t *)irwls_dev_alloc((size_t)(rows + 1) * sizeof(int64_t));
        dd->col = (int32_t *)irwls_dev_alloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
        dd->val = (double *)irwls_dev_alloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
        irwls_dev_h2d(dd->rowptr, rp, (size_t)(rows + 1) * sizeof(int64_t));
        j.what = 0;
        irwls_dev_h2d_chunked(dd->col, (size_t)nnz * sizeof(int32_t), kUploadChunk, fill_csr, &j);
        j.what = 1;
        irwls_dev_h2d_chunked(dd->val, (size_t)nnz * sizeof(double), kUploadChunk, fill_csr, &j);
        free(rp);
        dd->nnz_per_row = rows > 0 ? (double)nnz / rows : 0.0;
        dd->nrm = dd->qv;
    }
    dd->fp_yq = fp_labels(ds, rows);
    dd->fp_mean = fp_means(ds, rows);
    return dd;
}

static void destroy(irwls_devdata *dd) {
    if (dd->Xh) irwls_dev_free(dd->Xh);
    if (dd->Xl) irwls_dev_free(dd->Xl);
    if (dd->nrm32) irwls_dev_free(dd->nrm32);
    irwls_dev_free(dd->X);
    irwls_dev_free(dd->rowptr);
    irwls_dev_free(dd->col);
    irwls_dev_free(dd->val);
    if (dd->nrm != dd->qv) irwls_dev_free(dd->nrm);
    irwls_dev_free(dd->qv);
    irwls_dev_free(dd->y);
    free(dd);
}

/* A pinned copy is the dataset's when the arrays are the same and the
   fingerprints of its labels, norms and mean rows still match: labels or
   norms changed in place (one-vs-rest relabelling of a shared dataset) are
   copied again into the same device buffers, changed mean rows (whose entry
   count may differ) take a new upload.  Called with g_dd_mu held; the copies
   are complete before another thread can use the entry.  Feature values
   changed in place are not detected: unpin a dataset before editing them. */
static irwls_devdata *pinned_match(int slot, const svm_dataset *ds) {
    irwls_devdata *p = g_pinned[slot];
    if (p->key != ds->features || p->key_x != ds->x || p->key_y != ds->y || p->key_q != ds->quadratic_value ||
        p->l != ds->l || p->dense != (ds->sparse == 0))
        return NULL;
    if (fp_means(ds, p->rows) != p->fp_mean) {
        const int with_means = p->rows > p->l;
        const int base = p->base, d = p->d;
        destroy(p);
        p = upload(ds, with_means, base, d);
        irwls_dev_sync();
        p->pinned = 1;
        g_pinned[slot] = p;
        return p;
    }
    const uint64_t fy = fp_labels(ds, p->rows);
    if (fy != p->fp_yq) {
        const size_t rr = (size_t)(p->rows > 0 ? p->rows : 1);
        double *hy = (double *)irwls_xcalloc(rr, sizeof(double));
        for (int i = 0; i < p->rows; i++) hy[i] = ds->y ? ds->y[i] : 0.0;
        irwls_dev_h2d(p->y, hy, (size_t)p->rows * sizeof(double));
        if (ds->quadratic_value) irwls_dev_h2d(p->qv, ds->quadratic_value, (size_t)p->rows * sizeof(double));
        irwls_dev_sync();
        free(hy);
        p->fp_yq = fy;
    }
    return p;
}

irwls_devdata *irwls_devdata_acquire(const svm_dataset *ds, int with_means, int base, int min_cols) {
    irwls_dev_init();
    pthread_mutex_lock(&g_dd_mu);
    for (int i = 0; i < MAX_PINNED; i++) {
        irwls_devdata *p = g_pinned[i];
        if (p && p->key == ds->features && (p->rows >= ds->l + (with_means ? 2 : 0)) &&
            (base < 0 || p->base == base) && p->d >= min_cols) {
            p = pinned_match(i, ds);
            if (!p) continue;
            pthread_mutex_unlock(&g_dd_mu);
            return p;
        }
    }
    pthread_mutex_unlock(&g_dd_mu);
    /* an unpinned copy belongs to the calling thread alone */
    return upload(ds, with_means, base, min_cols);
}

int irwls_devdata_pinned_range(const svm_dataset *ds, int *base, int *maxindex) {
    int found = 0;
    pthread_mutex_lock(&g_dd_mu);
    for (int i = 0; i < MAX_PINNED && !found; i++) {
        const irwls_devdata *p = g_pinned[i];
        if (p && p->key == ds->features && p->key_x == ds->x && p->l == ds->l && p->dense == (ds->sparse == 0)) {
            *base = p->base;
            *maxindex = p->base + p->d - 1;
            found = 1;
        }
    }
    pthread_mutex_unlock(&g_dd_mu);
    return found;
}

void irwls_devdata_release(irwls_devdata *dd) {
    if (!dd || dd->pinned) return;
    destroy(dd);
}

irwls_rows irwls_devdata_rows(const irwls_devdata *dd, const int32_t *d_idx, int n) {
    irwls_rows r;
    memset(&r, 0, sizeof(r));
    if (dd->dense) {
        r.X = dd->X;
        r.ldx = dd->d;
    } else {
        r.rowptr = dd->rowptr;
        r.col = dd->col;
        r.val = dd->val;
        r.nnz_per_row = dd->nnz_per_row;
    }
    r.nrm = dd->nrm;
    r.idx = d_idx;
    r.n = n;
    return r;
}

static int use_bf16(const irwls_devdata *dd) {
    const int p = irwls_opts()->precision;
    return (p == 1 || p == 3) && dd->dense && dd->d > 0;
}

static irwls_rows_bf16 rows_bf16(irwls_devdata *dd, const int32_t *d_idx, int n) {
    pthread_mutex_lock(&g_dd_mu);
    if (!dd->Xh) {
        dd->ldb = (dd->d + 31) / 32 * 32;
        uint16_t *Xh = (uint16_t *)irwls_dev_alloc((size_t)dd->rows * dd->ldb * sizeof(uint16_t));
        uint16_t *Xl = (uint16_t *)irwls_dev_alloc((size_t)dd->rows * dd->ldb * sizeof(uint16_t));
        float *nrm32 = (float *)irwls_dev_alloc((size_t)dd->rows * sizeof(float));
        irwls_dev_bf16_split(dd->X, dd->d, dd->rows, dd->d, Xh, Xl, dd->ldb, nrm32);
        irwls_dev_sync();   /* complete before another thread's stream reads it */
        dd->Xl = Xl;
        dd->nrm32 = nrm32;
        dd->Xh = Xh;
    }
    pthread_mutex_unlock(&g_dd_mu);
    irwls_rows_bf16 r;
    r.Xh = dd->Xh;
    r.Xl = dd->Xl;
    r.ld = dd->ldb;
    r.nrm = dd->nrm32;
    r.idx = d_idx;
    r.n = n;
    return r;
}

void irwls_kmv(irwls_devdata *dd, const int32_t *d_ridx, int rn, const int32_t *d_cidx, int cn, int kernel,
               double gamma, const double *d_w, double a, double b, double cst, const double *d_rs, double *d_out) {
    if (use_bf16(dd)) {
        irwls_rows_bf16 r = rows_bf16(dd, d_ridx, rn), c = rows_bf16(dd, d_cidx, cn);
        irwls_dev_kernel_matvec_bf16(&r, &c, kernel, gamma, d_w, a, b, cst, d_rs, d_out);
    } else {
        irwls_rows r = irwls_devdata_rows(dd, d_ridx, rn), c = irwls_devdata_rows(dd, d_cidx, cn);
        irwls_dev_kernel_matvec(&r, &c, dd->d, kernel, gamma, d_w, a, b, cst, d_rs, d_out);
    }
}

void irwls_kmat(irwls_devdata *dd, const int32_t *d_ridx, int rn, const int32_t *d_cidx, int cn, int kernel,
                double gamma, double *d_K, int ldk) {
    if (use_bf16(dd)) {
        irwls_rows_bf16 r = rows_bf16(dd, d_ridx, rn), c = rows_bf16(dd, d_cidx, cn);
        irwls_dev_kernel_matrix_bf16(&r, &c, kernel, gamma, d_K, ldk);
    } else {
        irwls_rows r = irwls_devdata_rows(dd, d_ridx, rn), c = irwls_devdata_rows(dd, d_cidx, cn);
        irwls_dev_kernel_matrix(&r, &c, dd->d, kernel, gamma, d_K, ldk);
    }
}

void irwls_gpu_pin_dataset(svm_dataset ds, int with_means) {
    irwls_dev_init();
    pthread_mutex_lock(&g_dd_mu);
    int free_slot = -1;
    for (int i = 0; i < MAX_PINNED; i++) {
        if (g_pinned[i] && g_pinned[i]->key == ds.features) {
            /* the same arrays, with at least the rows asked for: keep it
               (refreshed); otherwise this pin replaces the old entry */
            if (g_pinned[i]->rows >= ds.l + (with_means ? 2 : 0) && pinned_match(i, &ds)) {
                pthread_mutex_unlock(&g_dd_mu);
                return;
            }
            destroy(g_pinned[i]);
            g_pinned[i] = NULL;
        }
        if (!g_pinned[i] && free_slot < 0) free_slot = i;
    }
    if (free_slot < 0) {
        fprintf(stderr, "libirwls: more than %d pinned datasets\n", MAX_PINNED);
        exit(2);
    }
    irwls_devdata *dd = upload(&ds, with_means, -1, 0);
    irwls_dev_sync();
    dd->pinned = 1;
    g_pinned[free_slot] = dd;
    pthread_mutex_unlock(&g_dd_mu);
}

void irwls_gpu_unpin_dataset(svm_dataset ds) {
    pthread_mutex_lock(&g_dd_mu);
    for (int i = 0; i < MAX_PINNED; i++) {
        if (g_pinned[i] && g_pinned[i]->key == ds.features) {
            destroy(g_pinned[i]);
            g_pinned[i] = NULL;
        }
    }
    pthread_mutex_unlock(&g_dd_mu);
}

void irwls_gpu_shutdown(void) {
    for (int i = 0; i < MAX_PINNED; i++) {
        if (g_pinned[i]) {
            destroy(g_pinned[i]);
            g_pinned[i] = NULL;
        }
    }
    irwls_mixed_release();
    irwls_dev_shutdown();
}

/* ---- working-set buffers ------------------------------------------------ */
void irwls_ws_reserve(irwls_ws *ws, int cap) {
    if (ws->cap >= cap) return;
    irwls_ws_free(ws);
    const size_t c = (size_t)(cap > 0 ? cap : 1);
    ws->cap = (int)c;
    ws->Kw = (double *)irwls_dev_alloc(c * c * sizeof(double));
    ws->H = (double *)irwls_dev_alloc((c + 2) * (c + 1) * sizeof(double));
    ws->esub = (double *)irwls_dev_alloc((c + 1) * sizeof(double));
    ws->G13 = (double *)irwls_dev_alloc((c + 1) * sizeof(double));
    ws->x = (double *)irwls_dev_alloc((c + 2) * sizeof(double));
    ws->yl = (double *)irwls_dev_alloc((c + 1) * sizeof(double));
    ws->gin = (double *)irwls_dev_alloc((c + 1) * sizeof(double));
    ws->beta = (double *)irwls_dev_alloc((c + 1) * sizeof(double));
    ws->bnew = (double *)irwls_dev_alloc((c + 1) * sizeof(double));
    ws->nc_cst = (double *)irwls_dev_alloc(2 * sizeof(double));
    const size_t ni = (c + 2) / 2, nd = c + 1;   /* c + 1 int32 or double entries, in doubles */
    ws->span_len[0] = ni + 3 * nd;
    ws->span_len[1] = ni + nd;
    ws->span_len[2] = 2 * ni + 2 * nd;
    ws->span_off[0] = 0;
    ws->span_off[1] = ws->span_len[0];
    ws->span_off[2] = ws->span_len[0] + ws->span_len[1];
    const size_t tot = ws->span_off[2] + ws->span_len[2];
    ws->pk = (double *)irwls_dev_alloc(tot * sizeof(double));
    ws->hpk = (double *)irwls_xcalloc(tot, sizeof(double));
    double *d = ws->pk, *h = ws->hpk;
    size_t o = 0;
    ws->S1 = (int32_t *)(d + o);
    ws->hS1 = (int32_t *)(h + o);
    o += ni;
    ws->ys = d + o;
    ws->hys = h + o;
    o += nd;
    ws->as = d + o;
    ws->has = h + o;
    o += nd;
    ws->rhs = d + o;
    ws->hrhs = h + o;
    o += nd;
    ws->cols = (int32_t *)(d + o);
    ws->hcols = (int32_t *)(h + o);
    o += ni;
    ws->w = d + o;
    ws->hw = h + o;
    o += nd;
    ws->S1g = (int32_t *)(d + o);
    ws->hS1g = (int32_t *)(h + o);
    o += ni;
    ws->S3 = (int32_t *)(d + o);
    ws->hS3 = (int32_t *)(h + o);
    o += ni;
    ws->wg = d + o;
    ws->hwg = h + o;
    o += nd;
    ws->rs = d + o;
    ws->hrs = h + o;
}

void irwls_ws_upload(irwls_ws *ws, int span) {
    irwls_dev_h2d(ws->pk + ws->span_off[span], ws->hpk + ws->span_off[span], ws->span_len[span] * sizeof(double));
}

void irwls_ws_free(irwls_ws *ws) {
    if (ws->cap == 0) return;
    irwls_dev_free(ws->Kw);
    irwls_dev_free(ws->H);
    irwls_dev_free(ws->esub);
    irwls_dev_free(ws->G13);
    irwls_dev_free(ws->x);
    irwls_dev_free(ws->yl);
    irwls_dev_free(ws->gin);
    irwls_dev_free(ws->beta);
    irwls_dev_free(ws->bnew);
    irwls_dev_free(ws->nc_cst);
    irwls_dev_free(ws->pk);
    free(ws->hpk);
    memset(ws, 0, sizeof(*ws));
}
