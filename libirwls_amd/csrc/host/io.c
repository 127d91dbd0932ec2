/*
 * Data model and I/O: dataset readers, in-memory builders, the binary model
 * file and the prediction text output.
 *
 * Semantics follow the reference readers and writers
 * (src/IOStructures.c:53-848): rows end with index -1, |x|^2 is a running sum
 * of squares in file order, `sparse` is set when the k-th index of a row
 * differs from the k-th entry of the feature array, and the labelled readers
 * append the positive-class and negative-class mean rows l and l+1 (the
 * negative mean's norm uses the positive count, src/IOStructures.c:259,477,
 * kept because budgeted models store that norm).  The model file is the
 * reference's raw host-endian dump, byte for byte.
 */
#define _POSIX_C_SOURCE 200809L
#include <ctype.h>
#include <limits.h>
#include <pthread.h>
#include <unistd.h>
#include <math.h>
#include <errno.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"

/* struct layouts every existing caller and model file relies on */
_Static_assert(sizeof(svm_sample) == 16, "svm_sample layout");
_Static_assert(offsetof(svm_sample, value) == 8, "svm_sample layout");
_Static_assert(sizeof(svm_dataset) == 48, "svm_dataset layout");
_Static_assert(sizeof(properties) == 72, "properties layout");
_Static_assert(sizeof(predictProperties) == 32, "predictProperties layout");
_Static_assert(sizeof(model) == 72, "model layout");

void freeDataset(svm_dataset data) {
    irwls_gpu_unpin_dataset(data);   /* a pinned device copy goes with the host arrays */
    free(data.y);
    free(data.quadratic_value);
    free(data.x);
    free(data.features);
}

void freeModel(model modelo) {
    free(modelo.weights);
    free(modelo.quadratic_value);
    free(modelo.x);
    free(modelo.features);
}

static void bad_format(void) {
    fprintf(stderr, "Wrong file format\n");
    exit(2);
}

/* Row store used while parsing: CSR with the reference's 1-based indices. */
typedef struct rowbuf {
    int64_t *rowptr;
    int32_t *idx;
    double *val;
    double *y;
    int n, ncap;
    int64_t nnz, zcap;
} rowbuf;

static void rb_init(rowbuf *b) {
    memset(b, 0, sizeof(*b));
    b->ncap = 1024;
    b->zcap = 16384;
    b->rowptr = (int64_t *)irwls_xmalloc((size_t)(b->ncap + 1) * sizeof(int64_t));
    b->y = (double *)irwls_xmalloc((size_t)b->ncap * sizeof(double));
    b->idx = (int32_t *)irwls_xmalloc((size_t)b->zcap * sizeof(int32_t));
    b->val = (double *)irwls_xmalloc((size_t)b->zcap * sizeof(double));
    b->rowptr[0] = 0;
}

static void rb_entry(rowbuf *b, int index, double v) {
    if (b->nnz == b->zcap) {
        b->zcap *= 2;
        b->idx = (int32_t *)realloc(b->idx, (size_t)b->zcap * sizeof(int32_t));
        b->val = (double *)realloc(b->val, (size_t)b->zcap * sizeof(double));
        if (!b->idx || !b->val) {
            fprintf(stderr, "libirwls: out of host memory\n");
            exit(2);
        }
    }
    b->idx[b->nnz] = index;
    b->val[b->nnz] = v;
    b->nnz++;
}

static void rb_endrow(rowbuf *b, double label) {
    if (b->n == b->ncap) {
        b->ncap *= 2;
        b->rowptr = (int64_t *)realloc(b->rowptr, (size_t)(b->ncap + 1) * sizeof(int64_t));
        b->y = (double *)realloc(b->y, (size_t)b->ncap * sizeof(double));
        if (!b->rowptr || !b->y) {
            fprintf(stderr, "libirwls: out of host memory\n");
            exit(2);
        }
    }
    b->y[b->n] = label;
    b->n++;
    b->rowptr[b->n] = b->nnz;
}

static void rb_free(rowbuf *b) {
    free(b->rowptr);
    free(b->idx);
    free(b->val);
    free(b->y);
}

/* Dataset from parsed rows with the reader semantics; with_means appends the
   two class-mean rows (labelled readers). */
static svm_dataset assemble(const int64_t *rowptr, const int32_t *idx, const double *val, const double *y, int n,
                            int with_means) {
    svm_dataset ds;
    memset(&ds, 0, sizeof(ds));
    int maxidx = 0;
    const int64_t nnz = rowptr[n];
    for (int64_t k = 0; k < nnz; k++)
        if (idx[k] > maxidx) maxidx = idx[k];
    const int extra = with_means ? 2 : 0;
    const int64_t elements = nnz + n + (with_means ? 2 * (int64_t)(maxidx + 2) : 0);
    ds.l = n;
    ds.y = (double *)irwls_xcalloc((size_t)n + extra, sizeof(double));
    ds.quadratic_value = (double *)irwls_xcalloc((size_t)n + extra, sizeof(double));
    ds.x = (svm_sample **)irwls_xcalloc((size_t)n + extra, sizeof(svm_sample *));
    ds.features = (svm_sample *)irwls_xcalloc((size_t)elements, sizeof(svm_sample));
    double *mp = with_means ? (double *)irwls_xcalloc((size_t)maxidx + 1, sizeof(double)) : NULL;
    double *mn = with_means ? (double *)irwls_xcalloc((size_t)maxidx + 1, sizeof(double)) : NULL;
    double cpos = 0.0, cneg = 0.0;
    int64_t j = 0;
    for (int i = 0; i < n; i++) {
        ds.x[i] = &ds.features[j];
        ds.y[i] = y ? y[i] : 0.0;
        if (ds.y[i] == 1.0) cpos += 1;
        else cneg += 1;
        int64_t pos = 0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++, pos++, j++) {
            svm_sample *f = &ds.features[j];
            f->index = idx[k];
            f->value = val[k];
            if (ds.features[pos].index != f->index) ds.sparse = 1;
            if (with_means) {
                if (ds.y[i] == 1.0) mp[f->index] += f->value;
                else mn[f->index] += f->value;
            }
            ds.quadratic_value[i] += f->value * f->value;
        }
        ds.features[j++].index = -1;
    }
    if (with_means) {
        ds.y[n] = 1.0;
        ds.x[n] = &ds.features[j];
        for (int f = 0; f <= maxidx; f++) {
            if (mp[f] == 0.0) continue;
            const double v = mp[f] / cpos;
            ds.features[j].index = f;
            ds.features[j].value = v;
            ds.quadratic_value[n] += v * v;
            j++;
        }
        ds.features[j++].index = -1;
        ds.y[n + 1] = -1.0;
        ds.x[n + 1] = &ds.features[j];
        for (int f = 0; f <= maxidx; f++) {
            if (mn[f] == 0.0) continue;
            const double skew = mn[f] / cpos;   /* norm uses the positive count */
            ds.features[j].index = f;
            ds.features[j].value = mn[f] / cneg;
            ds.quadratic_value[n + 1] += skew * skew;
            j++;
        }
        ds.features[j++].index = -1;
        free(mp);
        free(mn);
    }
    ds.maxdim = maxidx;
    return ds;
}

svm_dataset irwls_dataset_from_dense(const double *X, const double *y, int n, int d) {
    int64_t *rp = (int64_t *)irwls_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int32_t *ix = (int32_t *)irwls_xmalloc((size_t)n * d * sizeof(int32_t) + 4);
    for (int i = 0; i <= n; i++) rp[i] = (int64_t)i * d;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < d; k++) ix[(size_t)i * d + k] = k + 1;
    svm_dataset ds = assemble(rp, ix, X, y, n, 1);
    free(rp);
    free(ix);
    return ds;
}

svm_dataset irwls_dataset_from_csr(const long long *rowptr, const int *col, const double *val, const double *y,
                                   int n) {
    const int64_t nnz = (int64_t)rowptr[n];
    int64_t *rp = (int64_t *)irwls_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int32_t *ix = (int32_t *)irwls_xmalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    for (int i = 0; i <= n; i++) rp[i] = (int64_t)rowptr[i];
    for (int64_t k = 0; k < nnz; k++) {
        if (col[k] < 0 || col[k] >= INT_MAX - 1) {
            fprintf(stderr, "libirwls: irwls_dataset_from_csr: column %d of entry %lld is out of range\n", col[k],
                    (long long)k);
            exit(2);
        }
        ix[k] = col[k] + 1;
    }
    svm_dataset ds = assemble(rp, ix, val, y, n, 1);
    free(rp);
    free(ix);
    return ds;
}

/* Dataset from a dense row-major matrix with the conversion of the
   reference's Python extension (numpy2dataset and numpy2datasetWithAverage,
   python-package/pythonmodule.c:210-330), which differs from the readers':
   only nonzero values are stored (index k+1), |x|^2 is the sum of their
   pow(x, 2), maxdim is the column count and `sparse` is 1 as soon as any value
   is nonzero.  with_average (its budgeted trainer's conversion) appends rows
   l (y = +1) and l+1 (y = -1) holding the class means, accumulated in row
   order as x/count over the rows labelled exactly 1.0, resp. all others,
   with their "norms" stored as the plain sums of the mean values.  The
   extension writes the first mean row's terminator one slot too far, onto
   the second row's first entry (pythonmodule.c:293-294): the first mean row
   then lists indices 1..d followed by an (index 0, value 0) entry, and the
   second mean row is empty.  Both are reproduced, since the budgeted
   trainer's first two centroids are these rows. */
svm_dataset irwls_dataset_from_numpy(const double *X, const double *y, int n, int d, int with_average) {
    svm_dataset ds;
    memset(&ds, 0, sizeof(ds));
    const int extra = with_average ? 2 : 0;
    ds.l = n;
    ds.maxdim = d;
    ds.y = (double *)irwls_xcalloc((size_t)n + extra, sizeof(double));
    ds.quadratic_value = (double *)irwls_xcalloc((size_t)n + extra, sizeof(double));
    ds.x = (svm_sample **)irwls_xcalloc((size_t)n + extra, sizeof(svm_sample *));
    int64_t nz = 0;
    int pos = 0, neg = 0;
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < d; k++)
            if (X[(size_t)i * d + k] != 0.0) nz++;
        ds.y[i] = y ? y[i] : 0.0;
        if (ds.y[i] == 1.0)
            pos++;
        else
            neg++;
    }
    ds.sparse = nz > 0;
    const int64_t elements = nz + n;   /* the rows' entries and terminators */
    ds.features = (svm_sample *)irwls_xcalloc((size_t)elements + (with_average ? 2 * ((size_t)d + 1) : 0),
                                              sizeof(svm_sample));
    svm_sample *mp = NULL, *mn = NULL;
    if (with_average) {
        mp = ds.features + elements;
        mn = mp + d + 1;
        for (int k = 0; k < d; k++) {
            mp[k].index = k + 1;
            mn[k].index = k + 1;
        }
        mn[0].index = -1;   /* the misplaced terminator: mp[d] stays (0, 0.0), mn is empty */
        mn[d].index = -1;
        ds.y[n] = 1.0;
        ds.y[n + 1] = -1.0;
        ds.x[n] = mp;
        ds.x[n + 1] = mn;
    }
    int64_t j = 0;
    for (int i = 0; i < n; i++) {
        ds.x[i] = &ds.features[j];
        for (int k = 0; k < d; k++) {
            const double v = X[(size_t)i * d + k];
            if (v == 0.0) continue;
            ds.quadratic_value[i] += pow(v, 2.0);
            ds.features[j].index = k + 1;
            ds.features[j].value = v;
            j++;
            if (with_average) {
                if (ds.y[i] == 1.0)
                    mp[k].value += v / pos;
                else
                    mn[k].value += v / neg;
            }
        }
        ds.features[j++].index = -1;
    }
    if (with_average)
        for (int k = 0; k < d; k++) {
            ds.quadratic_value[n] += mp[k].value;
            ds.quadratic_value[n + 1] += mn[k].value;
        }
    return ds;
}

/* ---- text readers --------------------------------------------------------- */

static FILE *open_or_die(const char *filename) {
    if (filename == NULL) {
        fprintf(stderr, "File not specified");
        exit(2);
    }
    FILE *f = fopen(filename, "rb");
    if (!f) {
        fprintf(stderr, "File not found: %s\n", filename);
        exit(2);
    }
    return f;
}

static int blank_line(const char *s) {
    for (; *s; s++)
        if (!isspace((unsigned char)*s)) return 0;
    return 1;
}

/* One libsvm row "[label] idx:val idx:val ..." (the line without its
   newline) appended to b. */
static void parse_libsvm_line(char *line, int labelled, rowbuf *b) {
    char *p = line;
    double label = 0.0;
    if (labelled) {
        while (*p == ' ' || *p == '\t') p++;
        char *end;
        label = strtod(p, &end);
        if (end == p || !(*end == '\0' || isspace((unsigned char)*end))) bad_format();
        p = end;
    }
    int last = -1;
    for (;;) {
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '\0' || *p == '\n' || *p == '\r') break;
        char *end;
        errno = 0;
        long ix = strtol(p, &end, 10);
        if (end == p || *end != ':' || errno != 0 || ix <= last || ix >= INT_MAX) bad_format();
        p = end + 1;
        double v = strtod(p, &end);
        if (end == p || !(*end == '\0' || isspace((unsigned char)*end))) bad_format();
        p = end;
        rb_entry(b, (int)ix, v);
        last = (int)ix;
    }
    rb_endrow(b, label);
}

/* One CSV row "[label]<sep>v1<sep>v2..."; nonzero values are stored with
   their 1-based column position as index. */
static void parse_csv_line(char *line, const char *delim, int labelled, rowbuf *b) {
    char *save = NULL;
    char *tok = strtok_r(line, delim, &save);
    double label = 0.0;
    int column = 0;
    if (labelled) {
        if (!tok) bad_format();
        char *end;
        label = strtod(tok, &end);
        if (end == tok || !(*end == '\0' || isspace((unsigned char)*end))) bad_format();
        tok = strtok_r(NULL, delim, &save);
    }
    for (; tok; tok = strtok_r(NULL, delim, &save)) {
        column++;
        if (blank_line(tok)) break;
        char *end;
        double v = strtod(tok, &end);
        if (end == tok || !(*end == '\0' || isspace((unsigned char)*end))) bad_format();
        if (v != 0.0) rb_entry(b, column, v);
    }
    rb_endrow(b, label);
}

/* A run of whole lines [begin, end) of the file image, parsed by one thread
   into a row store of its own. */
typedef struct {
    char *begin, *end;
    int csv, labelled;
    const char *delim;
    rowbuf b;
} parse_chunk;

static void *parse_lines(void *arg) {
    parse_chunk *c = (parse_chunk *)arg;
    char *p = c->begin;
    while (p < c->end) {
        char *nl = (char *)memchr(p, '\n', (size_t)(c->end - p));
        char *eol = nl ? nl : c->end;
        *eol = '\0';   /* the image is NUL-terminated past its last byte */
        if (!blank_line(p)) {
            if (c->csv)
                parse_csv_line(p, c->delim, c->labelled, &c->b);
            else
                parse_libsvm_line(p, c->labelled, &c->b);
        }
        p = eol + 1;
    }
    return NULL;
}

/* Reads a whole text dataset: the file image is split at line boundaries
   into chunks of at least 256 KiB, parsed on up to 16 threads (one row store
   each), and the stores are joined in file order, so the result equals a
   single sequential pass line for line. */
static svm_dataset read_text(const char *filename, int csv, const char *sep, int labelled) {
    FILE *f = open_or_die(filename);
    if (fseek(f, 0, SEEK_END) != 0) bad_format();
    const long sz = ftell(f);
    if (sz < 0) bad_format();
    rewind(f);
    char *img = (char *)irwls_xmalloc((size_t)sz + 1);
    if ((long)fread(img, 1, (size_t)sz, f) != sz) {
        fprintf(stderr, "Error reading data file\n");
        exit(2);
    }
    img[sz] = '\0';
    fclose(f);
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    int nt = (int)(sz / (256L << 10));
    if (nt > 16) nt = 16;
    if (nt > ncpu) nt = (int)ncpu;
    if (nt < 1) nt = 1;
    parse_chunk *ch = (parse_chunk *)irwls_xcalloc((size_t)nt, sizeof(parse_chunk));
    char *pos = img;
    for (int t = 0; t < nt; t++) {
        char *end = img + sz;
        if (t + 1 < nt) {
            char *cut = img + (size_t)((double)sz * (t + 1) / nt);
            if (cut < pos) cut = pos;
            char *nl = (char *)memchr(cut, '\n', (size_t)(img + sz - cut));
            end = nl ? nl + 1 : img + sz;
        }
        ch[t].begin = pos;
        ch[t].end = end;
        ch[t].csv = csv;
        ch[t].labelled = labelled;
        ch[t].delim = (sep && *sep) ? sep : ",";
        rb_init(&ch[t].b);
        pos = end;
    }
    pthread_t *th = (pthread_t *)irwls_xcalloc((size_t)nt, sizeof(pthread_t));
    for (int t = 1; t < nt; t++)
        if (pthread_create(&th[t], NULL, parse_lines, &ch[t]) != 0) {
            fprintf(stderr, "libirwls: cannot start a reader thread\n");
            exit(2);
        }
    parse_lines(&ch[0]);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
    free(th);
    free(img);
    /* join the row stores in file order */
    int n = 0;
    int64_t nnz = 0;
    for (int t = 0; t < nt; t++) {
        n += ch[t].b.n;
        nnz += ch[t].b.nnz;
    }
    int64_t *rowptr = (int64_t *)irwls_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int32_t *idx = (int32_t *)irwls_xmalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    double *val = (double *)irwls_xmalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    double *y = (double *)irwls_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int r = 0;
    int64_t z = 0;
    rowptr[0] = 0;
    for (int t = 0; t < nt; t++) {
        const rowbuf *b = &ch[t].b;
        memcpy(idx + z, b->idx, (size_t)b->nnz * sizeof(int32_t));
        memcpy(val + z, b->val, (size_t)b->nnz * sizeof(double));
        memcpy(y + r, b->y, (size_t)b->n * sizeof(double));
        for (int i = 0; i < b->n; i++) rowptr[r + i + 1] = z + b->rowptr[i + 1];
        r += b->n;
        z += b->nnz;
        rb_free(&ch[t].b);
    }
    free(ch);
    svm_dataset ds = assemble(rowptr, idx, val, labelled ? y : NULL, n, labelled);
    free(rowptr);
    free(idx);
    free(val);
    free(y);
    return ds;
}

static svm_dataset read_libsvm(const char *filename, int labelled) { return read_text(filename, 0, NULL, labelled); }
static svm_dataset read_csv(const char *filename, const char *sep, int labelled) {
    return read_text(filename, 1, sep, labelled);
}

svm_dataset readTrainFile(char filename[]) { return read_libsvm(filename, 1); }
svm_dataset readTrainFileCSV(char filename[], char *separator) { return read_csv(filename, separator, 1); }
svm_dataset readUnlabeledFile(char filename[]) { return read_libsvm(filename, 0); }
svm_dataset readUnlabeledFileCSV(char filename[], char *separator) { return read_csv(filename, separator, 0); }

/* ---- binary dataset cache ---------------------------------------------------
   "IRWLSDS1", then int32 l, sparse, maxdim, rows (l, or l + 2 with the
   class-mean rows), int64 entry count, then y[rows] (zeros when the dataset
   has no labels), quadratic_value[rows] and the entries (16-byte index/value
   pairs, each row ended by index -1) in row order.  Row k starts after the
   k-th terminator, as in the model file. */
static const char kDsMagic[8] = {'I', 'R', 'W', 'L', 'S', 'D', 'S', '1'};

int irwls_dataset_save(const svm_dataset *ds, int with_means, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const int32_t head[4] = {ds->l, ds->sparse, ds->maxdim, ds->l + (with_means ? 2 : 0)};
    const int rows = head[3];
    int64_t ne = 0;
    for (int i = 0; i < rows; i++) {
        const svm_sample *s = ds->x[i];
        while (s->index != -1) s++;
        ne += (s - ds->x[i]) + 1;
    }
    int ok = fwrite(kDsMagic, 1, 8, f) == 8 && fwrite(head, sizeof(int32_t), 4, f) == 4 &&
             fwrite(&ne, sizeof(int64_t), 1, f) == 1;
    double *y = (double *)irwls_xcalloc((size_t)(rows > 0 ? rows : 1), sizeof(double));
    if (ds->y) memcpy(y, ds->y, (size_t)rows * sizeof(double));
    ok = ok && fwrite(y, sizeof(double), (size_t)rows, f) == (size_t)rows;
    ok = ok && fwrite(ds->quadratic_value, sizeof(double), (size_t)rows, f) == (size_t)rows;
    free(y);
    for (int i = 0; ok && i < rows; i++) {
        const svm_sample *s = ds->x[i];
        while (s->index != -1) s++;
        const size_t k = (size_t)(s - ds->x[i]) + 1;
        ok = fwrite(ds->x[i], sizeof(svm_sample), k, f) == k;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

svm_dataset irwls_dataset_load(const char *path) {
    FILE *f = open_or_die(path);
    char magic[8];
    int32_t head[4];
    int64_t ne = 0;
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kDsMagic, 8) != 0 || fread(head, sizeof(int32_t), 4, f) != 4 ||
        fread(&ne, sizeof(int64_t), 1, f) != 1 || head[0] < 0 || head[3] < head[0] || ne < head[3])
        bad_format();
    svm_dataset ds;
    memset(&ds, 0, sizeof(ds));
    ds.l = head[0];
    ds.sparse = head[1];
    ds.maxdim = head[2];
    const int rows = head[3];
    const size_t rr = (size_t)(rows > 0 ? rows : 1);
    ds.y = (double *)irwls_xmalloc(rr * sizeof(double));
    ds.quadratic_value = (double *)irwls_xmalloc(rr * sizeof(double));
    ds.x = (svm_sample **)irwls_xcalloc(rr, sizeof(svm_sample *));
    ds.features = (svm_sample *)irwls_xmalloc((size_t)(ne > 0 ? ne : 1) * sizeof(svm_sample));
    if (fread(ds.y, sizeof(double), (size_t)rows, f) != (size_t)rows ||
        fread(ds.quadratic_value, sizeof(double), (size_t)rows, f) != (size_t)rows ||
        fread(ds.features, sizeof(svm_sample), (size_t)ne, f) != (size_t)ne)
        bad_format();
    fclose(f);
    int row = 0;
    if (rows > 0) ds.x[row++] = ds.features;
    for (int64_t k = 0; k < ne; k++)
        if (ds.features[k].index == -1 && row < rows && k + 1 < ne) ds.x[row++] = &ds.features[k + 1];
    if (row != rows || (ne > 0 && ds.features[ne - 1].index != -1)) bad_format();
    return ds;
}

/* ---- model file ------------------------------------------------------------ */

void storeModel(model *mod, FILE *Output) {
    size_t w = 0;
    w += fwrite(&mod->Kgamma, sizeof(double), 1, Output);
    w += fwrite(&mod->bias, sizeof(double), 1, Output);
    w += fwrite(&mod->maxdim, sizeof(int), 1, Output);
    w += fwrite(&mod->kernelType, sizeof(int), 1, Output);
    w += fwrite(&mod->sparse, sizeof(int), 1, Output);
    w += fwrite(&mod->nSVs, sizeof(int), 1, Output);
    w += fwrite(&mod->nElem, sizeof(int), 1, Output);
    if (mod->nSVs > 0) {
        w += fwrite(mod->weights, sizeof(double), (size_t)mod->nSVs, Output);
        w += fwrite(mod->quadratic_value, sizeof(double), (size_t)mod->nSVs, Output);
    }
    if (mod->nElem > 0) w += fwrite(mod->features, sizeof(svm_sample), (size_t)mod->nElem, Output);
    (void)w;
    fflush(Output);
}

void readModel(model *mod, FILE *Input) {
    size_t r = 0;
    r += fread(&mod->Kgamma, sizeof(double), 1, Input);
    r += fread(&mod->bias, sizeof(double), 1, Input);
    r += fread(&mod->maxdim, sizeof(int), 1, Input);
    r += fread(&mod->kernelType, sizeof(int), 1, Input);
    r += fread(&mod->sparse, sizeof(int), 1, Input);
    r += fread(&mod->nSVs, sizeof(int), 1, Input);
    r += fread(&mod->nElem, sizeof(int), 1, Input);
    const int ns = mod->nSVs > 0 ? mod->nSVs : 0;
    const int ne = mod->nElem > 0 ? mod->nElem : 0;
    mod->weights = (double *)irwls_xmalloc((size_t)(ns ? ns : 1) * sizeof(double));
    mod->quadratic_value = (double *)irwls_xmalloc((size_t)(ns ? ns : 1) * sizeof(double));
    r += fread(mod->weights, sizeof(double), (size_t)ns, Input);
    r += fread(mod->quadratic_value, sizeof(double), (size_t)ns, Input);
    mod->x = (svm_sample **)irwls_xmalloc((size_t)(ns ? ns : 1) * sizeof(svm_sample *));
    mod->features = (svm_sample *)irwls_xcalloc((size_t)(ne ? ne : 1), sizeof(svm_sample));
    r += fread(mod->features, sizeof(svm_sample), (size_t)ne, Input);
    (void)r;
    /* row k starts after the k-th terminator (src/IOStructures.c:817-824) */
    if (ns > 0) mod->x[0] = &mod->features[0];
    int row = 1;
    for (int k = 0; k < ne; k++) {
        if (mod->features[k].index == -1) {
            if (row < ns) mod->x[row] = &mod->features[k + 1];
            row++;
        }
    }
}

void writeOutput(char fileoutput[], double *predictions, int size) {
    FILE *f = fopen(fileoutput, "w+");
    if (!f) {
        fprintf(stderr, "Cannot open output file: %s\n", fileoutput);
        exit(2);
    }
    for (int i = 0; i < size; i++) fprintf(f, "%lf\n", predictions[i]);
    fclose(f);
}
