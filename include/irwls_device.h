/*
 * Thin C-ABI over the hand-written CDNA4 (gfx950) HIP kernels.
 *
 * The C host code of the trainers (libirwls_amd/csrc/host, C99) drives the
 * GPU only through these functions: plain pointers and sizes, no HIP or
 * torch types.  Pointers named d_* are device pointers obtained from
 * irwls_dev_alloc.  All work is issued on one library-owned HIP stream, in
 * order; functions that return host data synchronise that stream.
 * HIP errors print a message to stderr and exit(2), the reference's
 * convention for fatal errors (src/IOStructures.c:98-107).
 *
 * This layer has no counterpart in the reference: it replaces the scalar
 * kernelFunction loops (src/kernels.c:63-118) and the BLAS/LAPACK calls of
 * src/ParallelAlgorithms.c and src/budgeted-train.c.
 */
#ifndef IRWLS_DEVICE_H
#define IRWLS_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- context, memory, copies ------------------------------------------ */
int   irwls_dev_init(void);                 /* idempotent; returns 0 */
void  irwls_dev_shutdown(void);
void *irwls_dev_alloc(size_t bytes);        /* zero-filled */
void  irwls_dev_free(void *d_ptr);          /* waits for the stream, keeps the block for reuse */
void  irwls_dev_h2d(void *d_dst, const void *src, size_t bytes);   /* synchronous */
/* Host-to-device copy of `total` bytes that the caller produces chunk by
   chunk: fill(arg, buf, off, len) writes bytes [off, off + len) of the
   source into the pinned buffer buf (len <= chunk); each chunk's transfer
   runs while the next is filled (two pinned buffers).  Returns when every
   chunk has been copied. */
void  irwls_dev_h2d_chunked(void *d_dst, size_t total, size_t chunk,
                            void (*fill)(void *arg, void *buf, size_t off, size_t len), void *arg);
void  irwls_dev_d2h(void *dst, const void *d_src, size_t bytes);   /* synchronous */
/* Two device-to-host copies with one wait for the stream (synchronous). */
void  irwls_dev_d2h2(void *dst1, const void *d_src1, size_t bytes1, void *dst2, const void *d_src2, size_t bytes2);
void  irwls_dev_d2d(void *d_dst, const void *d_src, size_t bytes); /* stream-ordered */
void  irwls_dev_memset0(void *d_dst, size_t bytes);                /* fill kernel */
/* Plain copy kernel (16-byte aligned pointers, bytes a multiple of 16): its
   traffic is known, which calibrates the hardware traffic counters. */
void  irwls_dev_copy_kernel(void *d_dst, const void *d_src, size_t bytes);
void  irwls_dev_sync(void);
void  irwls_dev_trim(void);                 /* release scratch and the cached free device blocks */

/* ---- per-kernel timing (HIP events on the library stream) -------------- */
enum irwls_timer_kind {
    IRWLS_T_KMV = 0,      /* fused kernel-matrix x vector (E-update, GIN, predict) */
    IRWLS_T_KMAT,         /* kernel-matrix materialisation (working-set block) */
    IRWLS_T_POTF2,        /* diagonal-block factorisation */
    IRWLS_T_TRSM,         /* panel solve */
    IRWLS_T_SYRK,         /* trailing update */
    IRWLS_T_POTRS,        /* triangular solves */
    IRWLS_T_KKT,          /* KKT matrix assembly */
    IRWLS_T_GEMV,         /* working-set gathers (inner error update, G13) */
    IRWLS_T_SGMA,         /* SGMA residual / gain */
    IRWLS_T_CSR,          /* CSR kernel columns */
    IRWLS_T_SYRKD,        /* IRWLSpar weighted Gram update */
    IRWLS_T_MATVEC,       /* IRWLSpar e = y - Phi beta */
    IRWLS_T_PCG,          /* mixed-precision refinement (preconditioned CG) */
    IRWLS_T_BF16,         /* BF16 hi/lo kernel tiles (mixed precision) */
    IRWLS_T_MLP,          /* proposal network training / inference */
    IRWLS_T_GAIN16,       /* SGMA gain of at most 16 candidates (HBM-bound pass over KSC) */
    IRWLS_T_COUNT
};
void irwls_dev_timing(int enable);          /* enable/disable and reset */
void irwls_dev_timing_mask(unsigned mask);  /* bit k set: kind k is timed (default: all) */
void irwls_dev_timing_stride(int stride);   /* time every stride-th launch of a kind (default 1: all) */
/* Accumulated launches, milliseconds, algorithmic flops and bytes. */
void irwls_dev_timer_get(int kind, long *launches, double *ms, double *flops, double *bytes);

/* ---- kernel tiles -------------------------------------------------------
 * Rows/columns of a kernel block are rows of a dense row-major FP64 matrix
 * X (leading dimension ldx) selected by an optional int32 index list
 * (NULL = identity), or rows of a CSR matrix.  k(x,z) is
 *   linear:  x.z
 *   RBF:     exp(-gamma * (|x|^2 + |z|^2 - 2 x.z))   (no clamping, as the
 *            reference's norm form, src/kernels.c:101-116)
 * with norms taken from the nrm vectors (indexed like the rows).          */
typedef struct irwls_rows {
    const double  *X;        /* dense rows (row-major), or NULL for CSR */
    int            ldx;
    const int64_t *rowptr;   /* CSR (0-based columns), used when X == NULL */
    const int32_t *col;
    const double  *val;
    const double  *nrm;      /* |x|^2 per row of X / CSR (before index mapping) */
    const int32_t *idx;      /* optional row selection */
    int            n;        /* number of selected rows */
    double         nnz_per_row; /* mean nonzeros per CSR row (traffic accounting only; 0 if unknown) */
} irwls_rows;

/* |x|^2 of n dense rows, summed sequentially in feature order. */
void irwls_dev_row_norms(const double *d_X, int ldx, int n, int d, double *d_nrm);

/* out[i] = a*out[i] + rs[i]*(b * sum_j w[j] k(r_i, c_j)) + c
   (rs NULL means 1).  Deterministic: fixed summation order for a given
   (rows, cols) shape.  With the linear kernel the block is not formed:
   u = sum_j w[j] c_j over the features, then out from r_i . u. */
void irwls_dev_kernel_matvec(const irwls_rows *r, const irwls_rows *c, int d,
                             int kernel, double gamma, const double *d_w,
                             double a, double b, double cst, const double *d_rs,
                             double *d_out);

/* irwls_dev_kernel_matvec without any device allocation.  It has no
   counterpart in the reference: it is for callers that cannot allocate
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same kernels and summation order as irwls_dev_kernel_matvec,
   so the same result bit for bit; the column splits' partial sums live in a
   fixed device array of IRWLS_NOALLOC_SCRATCH_BYTES per GPU (held until the
   process ends) instead of the library's scratch pool.  Columns must be
   dense (c->X set), and rows too for the RBF kernel.  Not timed.  Needs
   irwls_dev_init first.  All calls on one GPU share the array, so calls
   that overlap on one GPU (several host threads or streams) race.
   Returns 0, or -1 with a message on stderr when the inputs are CSR or the
   partial sums would exceed the array (nsplit x r->n doubles for RBF,
   (min(c->n, 64) + 1) x d for the linear kernel). */
#define IRWLS_NOALLOC_SCRATCH_BYTES (16 << 20)
int  irwls_dev_kernel_matvec_noalloc(const irwls_rows *r, const irwls_rows *c, int d,
                                     int kernel, double gamma, const double *d_w,
                                     double a, double b, double cst, const double *d_rs,
                                     double *d_out);

/* K[i*ldk + j] = k(r_i, c_j) for all i < r->n, j < c->n. */
void irwls_dev_kernel_matrix(const irwls_rows *r, const irwls_rows *c, int d,
                             int kernel, double gamma, double *d_K, int ldk);

/* A column side of CSR rows that only grows by appending rows (SGMA's
   centroids): its transposed column block, norms and feature masks live
   between kernel-block calls and are extended by the appended rows only,
   instead of being rebuilt for every call.  No counterpart in the reference
   (which evaluates kernelFunction pairwise); for the budgeted trainer.
   create returns NULL when cap x d doubles exceed 1 GiB (callers then use
   irwls_dev_kernel_matrix). */
typedef struct irwls_coltab irwls_coltab;
irwls_coltab *irwls_dev_coltab_create(int d, int cap);
void irwls_dev_coltab_destroy(irwls_coltab *t);
/* Columns held. */
int irwls_dev_coltab_size(const irwls_coltab *t);
/* Appends rows size(t) .. c->n - 1 of the CSR side c (c->idx selecting);
   its first size(t) rows must be the ones already held. */
void irwls_dev_coltab_extend(irwls_coltab *t, const irwls_rows *c);
/* irwls_dev_kernel_matrix(r, c) for the CSR rows r and the held columns:
   the same values bit for bit. */
void irwls_dev_kernel_matrix_coltab(const irwls_rows *r, const irwls_coltab *t, int kernel, double gamma,
                                    double *d_K, int ldk);

/* ---- full-training working-set helpers ---------------------------------- */
/* H (column-major N x N, leading dimension ldh, N = nS1+1):
   H[i][j] = Kw[s_i][s_j] y_i y_j + (i==j)/a_i, border y, corner 0.
   s = d_S1 (indices into Kw); d_y, d_a hold y and a already gathered by s. */
void irwls_dev_build_kkt(const double *d_Kw, int ldk, const int32_t *d_S1, int nS1,
                         const double *d_y, const double *d_a, double *d_H, int ldh);
/* out[i] = a*out[i] + rs[i] * (b * sum_j Kw[ri][cj] w[j]) + cst,
   ri = rows ? rows[i] : i, cj = cols ? cols[j] : j */
void irwls_dev_gemv_gather(const double *d_Kw, int ldk, const int32_t *d_rows, int nr,
                           const int32_t *d_cols, int nc, const double *d_w,
                           double a, double b, double cst, const double *d_rs,
                           double *d_out);

/* KKT right-hand side (src/full-train.c:160-168): rhs[i] = 1 - G13[i] -
   gin[S1[i]] for i < nS1 (G13 read as zero when use_g13 is 0), rhs[nS1] =
   last. */
void irwls_dev_kkt_rhs(const int32_t *d_S1, int nS1, const double *d_G13, int use_g13, const double *d_gin,
                       double last, double *d_rhs);
/* Weights after a KKT solve and the operands of the inner error update
   (src/full-train.c:205-249): bnew = 0, bnew[S1[i]] = x[i] y[S1[i]],
   bnew[S3[o]] = C y[S3[o]], bnew[l] = x[nS1]; the columns j < l with
   bnew[j] != beta[j] in ascending order, w = bnew[j] - beta[j], their count
   (int32 at word 0 of d_nc_cst) and -(bnew[l] - beta[l]) (word 1); then
   beta = bnew. */
void irwls_dev_inner_update(const double *d_x, const int32_t *d_S1, int nS1, const int32_t *d_S3, int nS3,
                            const double *d_y, double C, int l, double *d_beta, double *d_bnew, int32_t *d_cols,
                            double *d_w, double *d_nc_cst);
/* irwls_dev_gemv_gather over rows 0..nr-1 with the column count and the
   constant read from device memory (as irwls_dev_inner_update leaves them). */
void irwls_dev_gemv_gather_dn(const double *d_Kw, int ldk, int nr, const int32_t *d_cols, const double *d_w,
                              const double *d_nc_cst, double a, double b, double *d_out);

/* y = A x (trans 0) or A^T x (trans 1), A column-major m x n. */
void irwls_dev_gemv(const double *d_A, int m, int n, int lda, int trans, const double *d_x, double *d_y);

/* ---- dense Cholesky (column-major lower, LAPACK failure semantics) ------ */
/* Factor A (n x n, lda) in place.  Returns info: 0, or j+1 when pivot j is
   not positive (the Schur value is left at A(j,j) and the rest of the matrix
   is left as LAPACK's blocked potrf leaves it). */
int  irwls_dev_potrf(double *d_A, int n, int lda);
/* Solve L L^T x = b in place (b on device). */
void irwls_dev_potrs(const double *d_L, int n, int lda, double *d_b);
/* Factor A (n x n, lda >= n+1) and solve A x = rhs in one pass: rhs is
   placed in row n of A (the forward sweep rides along with the
   factorization), then one backward sweep gives x.  Failure semantics as
   irwls_dev_potrf (the solve uses the factor as it stands).  Returns info. */
int  irwls_dev_posv_aug(double *d_A, int n, int lda, const double *d_rhs, double *d_x);

/* Learned SGMA proposal: a 32 -> 128 -> 128 -> 1 MLP (ReLU).  Parameters
   are IRWLS_MLP_NPARAMS floats: W1[32][128], b1[128], W2[128][128], b2[128],
   w3[128], b3.  Loss: mean squared error over a batch. */
#define IRWLS_MLP_NPARAMS 20865
int  irwls_mlp_nparams(void);
int  irwls_mlp_npacked(void);
/* Gradient (grad, NPARAMS floats) and loss (1 float) of nb samples X [nb][32], targets T [nb]. */
void irwls_dev_mlp_grad(const float *d_P, const float *d_X, const float *d_T, int nb, float *d_grad, float *d_loss);
/* Adam step `step` (from 1), beta1 0.9, beta2 0.999, eps 1e-8. */
void irwls_dev_mlp_adam(float *d_P, const float *d_grad, float *d_m, float *d_v, int step, float lr);
void irwls_dev_mlp_forward_f32(const float *d_P, const float *d_X, int nb, float *d_Y);
/* BF16 inference: pack the weight matrices once (irwls_mlp_npacked() uint16), then score. */
void irwls_dev_mlp_pack_bf16(const float *d_P, uint16_t *d_Q);
void irwls_dev_mlp_infer_bf16(const uint16_t *d_Q, const float *d_P, const float *d_X, int nb, float *d_Y);
/* 32 features per candidate of a pool of P <= 1024 into F [P][32].  The
   positive fraction of the basis is pos_frac, or (float)*d_npos / (float)s
   when d_npos (a device count of positive centroids) is given. */
void irwls_dev_sgma_features(int P, int s, int B, const double *d_KNC, int ldk, const double *d_eta,
                             const double *d_gsub, const double *d_gsub2, const double *d_ycand,
                             const double *d_qvcand, float pos_frac, const int32_t *d_npos, float *d_F);
/* dst[j] = src[idx[j]], j < m */
void irwls_dev_gather_idx(const double *d_src, const int32_t *d_idx, int m, double *d_dst);

/* Mixed precision: dense rows as BF16 hi/lo pairs (x = hi + lo to about 16
   bits), row-major with ld a multiple of 32 (zero padded), and FP32 squared
   norms sum(hi^2 + 2 hi lo). */
typedef struct irwls_rows_bf16 {
    const uint16_t *Xh, *Xl;
    int ld;
    const float *nrm;      /* by original row */
    const int32_t *idx;    /* row selection, may be NULL */
    int n;
} irwls_rows_bf16;

void irwls_dev_bf16_split(const double *d_X, int ldx, int n, int d, uint16_t *d_Xh, uint16_t *d_Xl, int ldb,
                          float *d_nrm);
/* As irwls_dev_kernel_matvec / irwls_dev_kernel_matrix, dot products from
   three BF16 MFMA products with FP32 accumulation (dense rows only). */
void irwls_dev_kernel_matvec_bf16(const irwls_rows_bf16 *r, const irwls_rows_bf16 *c, int kernel, double gamma,
                                  const double *d_w, double a, double b, double cst, const double *d_rs,
                                  double *d_out);
void irwls_dev_kernel_matrix_bf16(const irwls_rows_bf16 *r, const irwls_rows_bf16 *c, int kernel, double gamma,
                                  double *d_K, int ldk);

/* Worker threads: irwls_dev_thread_begin gives the calling host thread a
   stream of its own (and its own scratch, staging and solver buffers), so
   several threads can drive the device at once; irwls_dev_thread_end waits
   for that stream and releases them.  Timers only record on the main
   thread's stream. */
void irwls_dev_thread_begin(void);
void irwls_dev_thread_end(void);
/* Creates worker streams on the calling thread until the pool holds n (the
   batched trainer calls it before starting its workers). */
void irwls_dev_reserve_streams(int n);
/* Seconds the calling thread has spent waiting for its stream (read-backs
   and synchronisations) since its last reset. */
double irwls_dev_wait_seconds(int reset);
void irwls_dev_release_thread_buffers(void);
void irwls_dev_release_cholesky_buffers(void);
void irwls_dev_release_pcg_buffers(void);
void irwls_dev_release_mlp_buffers(void);

/* Batched irwls_dev_posv_aug: nprob systems (augmented matrices A[p] with
   leading dimension lda[p] >= n[p]+1, right-hand sides rhs[p], solutions
   x[p], all device) factored and solved together, each with the
   arithmetic of the single-system call; info[p] (may be NULL) receives the
   factorization info.  The launch sequence is captured per batch class
   (count rounded up to a multiple of 16, largest size rounded up to 256)
   and replayed; its tile shape (64, 128 or 256 panel rows per work-group) is
   picked on the first call with a given rounded largest size by timing each
   on SPD systems of that batch's sizes.  Returns 0. */
int  irwls_dev_posv_aug_batched(int nprob, double *const *d_A, const int *n, const int *lda,
                                const double *const *d_rhs, double *const *d_x, int *info);
/* The same on round slot `round` (0..3, irwls_dev_posv_aug_batched uses 0):
   each slot has its own captured sequences and device buffers, so calls
   with different slots may run at the same time from different host threads
   (each on its own stream, irwls_dev_thread_begin); calls with one slot must
   not overlap. */
int  irwls_dev_posv_aug_batched_slot(int round, int nprob, double *const *d_A, const int *n, const int *lda,
                                     const double *const *d_rhs, double *const *d_x, int *info);
void irwls_dev_batched_autotune(int enable);

/* One inner IRWLS iteration's device work on a working set
   (src/full-train.c:158-249), all device pointers: the right-hand side
   rhs[i] = 1 - G13[i] - gin[S1[i]] (G13 read as zero unless use_g13),
   rhs[N-1] = last; the KKT matrix H (N = nS1 + 1, leading dimension
   ldh >= N+1) from the kernel block Kw (l x l, leading dimension ldk), ys
   and as (y and a over S1); its solve (LAPACK failure semantics) into x;
   then the weights after it and the error update e -= Kw(:, changed) dw +
   const, as irwls_dev_inner_update and irwls_dev_gemv_gather_dn do them
   (beta, bnew: l+1 entries; cols, w: l; nc_cst: 2 words). */
typedef struct irwls_inner_work {
    double *H;
    int N, ldh;
    const double *Kw;
    int ldk, l;
    const int32_t *S1;
    const double *ys, *as, *G13;
    int use_g13;
    const double *gin;
    double last;
    double *rhs, *x;
    const int32_t *S3;
    int nS3;
    const double *yl;
    double C;
    double *beta, *bnew;
    int32_t *cols;
    double *w, *nc_cst, *e;
} irwls_inner_work;
/* The inner iterations of nprob problems as one batched round on round
   slot `round` (as irwls_dev_posv_aug_batched_slot): every problem's
   arithmetic is that of the single-problem kernels. */
int  irwls_dev_inner_batched_slot(int round, int nprob, const irwls_inner_work *work);
/* Rows per work-group (in 64-row tiles) chosen for batches whose largest
   size rounds up to the same multiple of 256 as nmax, 0 if none yet; ms3
   (may be NULL) receives the three timings.  nprob is not used. */
int  irwls_dev_batched_choice(int nprob, int nmax, double *ms3);

/* Mixed precision.  FP32 Cholesky of the lower triangle of the FP64 matrix A
   (n x n, leading dimension lda), FP64 diagonal tiles, with the identity
   appended as rows n..2n-1 of F (FP32, column-major, ldf >= 2n): on return
   F holds L in rows 0..n-1 and Z = L^-T in rows n..2n-1, and ZT (n x n,
   ldz >= n) holds Z^T.  Returns the potrf info (0: positive definite). */
int  irwls_dev_potrf_inv_f32(const double *d_A, int n, int lda, float *d_F, int ldf, float *d_ZT, int ldz);
/* Preconditioned CG in FP64 for A X = B with two right-hand sides (B, X:
   device [2][n]); preconditioner Z Z^T from irwls_dev_potrf_inv_f32 (of this
   or an earlier matrix).  Stops when |b_c - A x_c| <= tol |b_c| for both
   columns (relres[c] on return).  Returns the iteration count, negated when
   maxit was reached without convergence. */
int  irwls_dev_pcg_spd2(const double *d_A, int n, int lda, const double *d_B, const float *d_F, int ldf,
                        const float *d_ZT, int ldz, double tol, int maxit, double *d_X, double *relres);
/* The same with a factor over another index set of size nf (ldf >= 2 nf,
   ldz >= nf): d_map[i] = factor position of row i or -1 (n entries),
   d_inv[f] = row at factor position f or -1 (nf entries).  Rows without a
   factor position are preconditioned by their diagonal. */
int  irwls_dev_pcg_spd2_mapped(const double *d_A, int n, int lda, const double *d_B, const float *d_F, int ldf,
                               const float *d_ZT, int ldz, int nf, const int32_t *d_map, const int32_t *d_inv,
                               double tol, int maxit, double *d_X, double *relres);
/* The KKT system's failed-last-pivot result from its two SPD solves
   (d_X = [u; z], u = A^-1 r1, z = A^-1 y; d_ys = y, n entries): s = -y^T z,
   b = (last - y^T u) / s^2, d_x = u - z b (n entries) and d_x[n] = b, as
   LAPACK's dpotrf/dpotrs leave it (src/ParallelAlgorithms.c:277-279, 441-445). */
void irwls_dev_kkt_closed_form(int n, const double *d_X, const double *d_ys, double last, double *d_x);
/* Rows removed from the factored set: with B = Z Z^T and R the factor
   positions of rs <= 256 removed rows (device), irwls_dev_pcg_schur_setup
   computes d_U = B(:, R) (nf x rs, column-major, leading dimension nf) and
   d_Sinv = S^-1 (rs x rs, column-major) for S = B(R, R), factored in FP64.
   With check set it waits for the device and returns that factorization's
   info (0: usable); without, it returns 0 at once (a solve with an unusable
   correction misses its residual target and reports it, as any solve that
   does not converge).
   irwls_dev_pcg_spd2_schur is irwls_dev_pcg_spd2_mapped whose kept rows are
   preconditioned by B_KK - B(K, R) S^-1 B(R, K), the exact inverse of the
   kept block of the factored matrix (rs = 0: the mapped form). */
int  irwls_dev_pcg_schur_setup(const float *d_ZT, int ldz, int nf, const int32_t *d_R, int rs, double *d_U,
                               double *d_Sinv, int check);
int  irwls_dev_pcg_spd2_schur(const double *d_A, int n, int lda, const double *d_B, const float *d_F, int ldf,
                              const float *d_ZT, int ldz, int nf, const int32_t *d_map, const int32_t *d_inv,
                              const int32_t *d_R, int rs, const double *d_U, const double *d_Sinv, double tol,
                              int maxit, double *d_X, double *relres);
/* Factor one SPD tile (n <= 64, column-major at d_A with leading dimension
   lda) in place and write L^-1 (column-major, leading dimension 64) and its
   transpose to d_Dinv (2 x 64 x 64).  With check set, waits and returns the
   potrf info; without, returns 0 at once. */
int  irwls_dev_tile_potrf_inv(double *d_A, int lda, int n, double *d_Dinv, int check);
/* FP64 factor of [A; I] (2n x n, column-major, ldg >= 2n; the caller writes
   the identity in rows n..2n-1): rows 0..n-1 receive L, rows n..2n-1
   Z = L^-T (upper triangular).  check as above. */
int  irwls_dev_potrf_ident(double *d_G, int n, int ldg, int check);

/* ---- budgeted training (SGMA, parametric IRWLS) ------------------------- */
/* C = alpha op(A) op(B) + beta C, all column-major (op = transpose when set). */
void irwls_dev_gemm(int transA, int transB, int m, int n, int k, double alpha, const double *d_A, int lda,
                    const double *d_B, int ldb, double beta, double *d_C, int ldc);
/* q[c] = sum_i (KSM[i*ldm + c] - sum_p KSC[i + p*n] Z[p + c*ldz])^2, c < ncand:
   the SGMA residual of every candidate column and its squared norm. */
void irwls_dev_sgma_gain(int n, int s, int ncand, const double *d_KSC, const double *d_Z, int ldz,
                         const double *d_KSM, int ldm, double *d_q);
/* eta[c] = 1 - sum_{e<s} KNC[c*ldk + e] Z[e + c*ldz] */
/* As irwls_dev_sgma_gain, plus qhalf = the same sums over the first half of
   the kernel's row tiles (64 rows each for ncand <= 16, else 128). */
void irwls_dev_sgma_gain_half(int n, int s, int ncand, const double *d_KSC, const double *d_Z, int ldz,
                              const double *d_KSM, int ldm, double *d_q, double *d_qhalf);
void irwls_dev_sgma_eta(int s, int ncand, const double *d_KNC, int ldk, const double *d_Z, int ldz, double *d_eta);
/* dst[i] = src[i*ld + col] */
void irwls_dev_gather_col(const double *d_src, int ld, int col, int n, double *d_dst);
/* M[c*ld + row] = v[c], c < count */
void irwls_dev_scatter_row(double *d_M, int ld, int row, int count, const double *d_v);
/* SGMA's factor growth (src/budgeted-train.c:247-279) on the device: with
   L2 = L^-1 k (s entries), IL2 = L2^T L^-1 (s entries) and kcc = k(c, c) in
   device memory, L3 = sqrt(kcc + 1e-5 - sum_i L2_i^2) (index order, no fused
   multiply-add), IL3 = 1/L3; writes row s of the column-major L and L^-1
   (leading dimension ld): (L2, L3) and (-IL3 IL2, IL3).  No host round trip. */
void irwls_dev_sgma_lrow(int s, int ld, const double *d_kcc, const double *d_L2, const double *d_IL2, double *d_L,
                         double *d_Li);
/* SGMA's choice of the step (src/budgeted-train.c:208-216) on the device:
   the first maximum of g_i = (eta_i > 0 ? (1/eta_i) q_i : 0) over nc
   candidates; st[0] = its position, *d_value = g, cent[s] = cand[position],
   and st[1] counts the chosen rows of label > 0 (row n, the positive class
   mean, counts; labels y on the device). */
void irwls_dev_sgma_pick(int nc, const double *d_q, const double *d_eta, const int32_t *d_cand, int s, int n,
                         const double *d_y, int32_t *d_cent, int32_t *d_st, double *d_value);
/* dst[i] = src[i * ld + *d_col], i < n (the column a device index names). */
void irwls_dev_gather_col_at(const double *d_src, int ld, const int32_t *d_col, int n, double *d_dst);
/* dst[j] = src[*d_row * ld + j], j < count. */
void irwls_dev_copy_row_at(const double *d_src, int ld, const int32_t *d_row, int count, double *d_dst);
/* The k highest of P <= 1024 scores, ties to the earlier index, one at a
   time as a serial scan with strict > picks them (a NaN in the first
   untaken place is picked, NaNs elsewhere never): out[r] = cidx[pick r]. */
void irwls_dev_sgma_topk(int P, const float *d_score, int k, const int32_t *d_cidx, int32_t *d_out);
/* The same, also writing the picks' positions: d_pos[r] = pick r. */
void irwls_dev_sgma_topk_pos(int P, const float *d_score, int k, const int32_t *d_cidx, int32_t *d_out,
                             int32_t *d_pos);
/* Shortlist operands gathered from the pool's (the learned proposal scores
   a subset of candidates whose KNC rows, Z columns and eta the pool step
   already computed): KNC2 row r and Z2 column r (s <= ld entries, leading
   dimension ld) and eta2[r] are those of pool position d_pos[r], r < k. */
void irwls_dev_sgma_gather_shortlist(int k, const int32_t *d_pos, int s, int ld, const double *d_KNC,
                                     const double *d_Z, const double *d_eta, double *d_KNC2, double *d_Z2,
                                     double *d_eta2);
/* K1 (ld ldk) = KC + sum_q (sw_q Phi[a_q])^T (sw_q Phi[a_q]) over the nact
   active rows a (NULL: rows 0..nact-1) of the row-major n x B matrix Phi;
   sw NULL means weight 1. */
void irwls_dev_wgram(int B, const double *d_Phi, const int32_t *d_act, const double *d_sw, int nact,
                     const double *d_KC, double *d_K1, int ldk);
/* K2[r] = sum_q (sw_q Phi[a_q][r]) (sw_q ymul y[a_q]) */
void irwls_dev_wvec(int B, const double *d_Phi, const int32_t *d_act, const double *d_sw, const double *d_y,
                    double ymul, int nact, double *d_K2);
/* e = y - Phi beta */
void irwls_dev_resid_rows(int n, int B, const double *d_Phi, const double *d_beta, const double *d_y, double *d_e);
/* D_i = C/(y_i e_i) clipped to M (0 when y_i e_i < 0); active rows (D != 0)
   compacted in row order into act with sw = sqrt(D).  Returns their count. */
int  irwls_dev_irls_weights(int n, const double *d_e, const double *d_y, double C, double M, double *d_D,
                            int32_t *d_act, double *d_sw);

#ifdef __cplusplus
}
#endif

#endif
