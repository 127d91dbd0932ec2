/*
 * Mixed-precision solve of subIRWLS's KKT system (src/full-train.c:169-199).
 *
 * The reference factors the indefinite matrix H = [A, y; y^T, 0] with
 * A = K_S1 o yy^T + diag(1/a) by LAPACK dpotrf, which stops at the last
 * pivot with the Schur value s = -y^T A^-1 y left on the diagonal, and dpotrs
 * then solves with that factor (src/ParallelAlgorithms.c:277-279, 441-445).
 * Its result is, in closed form,
 *     u = A^-1 r1,  z = A^-1 y,  s = -y^T z,
 *     b = (r_n - y^T u) / s^2,   beta = u - z b,
 * which only needs two solves with the SPD block A.  Here those run as
 * preconditioned CG in FP64 on the FP64 A (still intact in H) with the FP32
 * Cholesky factor of A (FP64 diagonal tiles, irwls_dev_potrf_inv_f32) as
 * preconditioner, to the relative residual irwls_gpu_options.refine_tol.
 *
 * Between inner iterations of one subIRWLS call, A changes in its diagonal
 * and its free set S1 mostly loses rows (rarely gains one): on the
 * n = 200000 workload a working set's free set shrinks by about 1400, 1000,
 * 500, 200, 40, 10, ... rows in its successive iterations.  The factor of the
 * earlier matrix then stays the preconditioner: rows still present map to
 * their factor positions and new ones take a diagonal step.  Up to 64
 * removed rows are corrected exactly: with B = Z Z^T the inverse of the old
 * matrix, B_KK - B_KR (B_RR)^-1 B_RK is the inverse of its kept block
 * (irwls_dev_pcg_schur_setup / irwls_dev_pcg_spd2_schur), so only the
 * diagonal change is left to CG (6-8 iterations against 2 after a fresh
 * factor).  Without the correction each removed row adds a rank-one
 * difference, about one more iteration each, so larger removals refactored.
 * More removed rows, many added ones, or CG missing its budget, refactor.
 * Where the diagonal's change dominates (a corrected solve needing more
 * than SCHUR_ITS iterations), the correction stops for that working set.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define SCHUR_MAX 64   /* removed rows the preconditioner corrects exactly */
#define SCHUR_ITS 10   /* a corrected solve needing more CG iterations stops the correction for its generation */

typedef struct {
    float *F;           /* 2 cap x cap: L over the identity rows' L^-T */
    float *ZT;          /* cap x cap */
    double *B, *X;      /* [2][cap] right-hand sides and solutions */
    int32_t *dmap, *dinv;
    int cap;
    int valid, n;       /* factor of the system of size n is held */
    long gen;           /* subIRWLS call it belongs to */
    int32_t *S1;        /* its free set */
    int32_t *map, *inv; /* host copies */
    int32_t *mark;      /* working-set index -> factor position + 1 */
    int mark_cap;
    int removed, added; /* of the last map_to_factor */
    long schur_off;     /* generation whose corrected solves took too many iterations */
    int32_t hR[SCHUR_MAX];   /* factor positions of the removed rows */
    int32_t *dR;        /* device copy */
    double *dU, *dSinv; /* B(:, R) (cap x SCHUR_MAX) and B(R, R)^-1 */
} mixed_cache;

/* per thread: the batched trainer runs one problem per host thread */
static __thread mixed_cache g_mc;
static __thread long g_gen;

long irwls_mixed_new_generation(void) { return ++g_gen; }

static void mc_reserve(int n) {
    if (g_mc.cap >= n) return;
    if (g_mc.F) irwls_dev_free(g_mc.F);
    if (g_mc.ZT) irwls_dev_free(g_mc.ZT);
    if (g_mc.B) irwls_dev_free(g_mc.B);
    if (g_mc.X) irwls_dev_free(g_mc.X);
    if (g_mc.dmap) irwls_dev_free(g_mc.dmap);
    if (g_mc.dinv) irwls_dev_free(g_mc.dinv);
    if (g_mc.dU) irwls_dev_free(g_mc.dU);
    free(g_mc.S1);
    free(g_mc.map);
    free(g_mc.inv);
    g_mc.dU = (double *)irwls_dev_alloc((size_t)n * SCHUR_MAX * sizeof(double));
    if (!g_mc.dR) g_mc.dR = (int32_t *)irwls_dev_alloc(SCHUR_MAX * sizeof(int32_t));
    if (!g_mc.dSinv) g_mc.dSinv = (double *)irwls_dev_alloc((size_t)SCHUR_MAX * SCHUR_MAX * sizeof(double));
    g_mc.F = (float *)irwls_dev_alloc((size_t)2 * n * n * sizeof(float));
    g_mc.dmap = (int32_t *)irwls_dev_alloc((size_t)n * sizeof(int32_t));
    g_mc.dinv = (int32_t *)irwls_dev_alloc((size_t)n * sizeof(int32_t));
    g_mc.map = (int32_t *)irwls_xmalloc((size_t)n * sizeof(int32_t));
    g_mc.inv = (int32_t *)irwls_xmalloc((size_t)n * sizeof(int32_t));
    g_mc.ZT = (float *)irwls_dev_alloc((size_t)n * n * sizeof(float));
    g_mc.B = (double *)irwls_dev_alloc((size_t)2 * n * sizeof(double));
    g_mc.X = (double *)irwls_dev_alloc((size_t)2 * n * sizeof(double));
    g_mc.S1 = (int32_t *)irwls_xmalloc((size_t)n * sizeof(int32_t));
    g_mc.cap = n;
    g_mc.valid = 0;
}

void irwls_mixed_release(void) {
    if (g_mc.F) irwls_dev_free(g_mc.F);
    if (g_mc.ZT) irwls_dev_free(g_mc.ZT);
    if (g_mc.B) irwls_dev_free(g_mc.B);
    if (g_mc.X) irwls_dev_free(g_mc.X);
    if (g_mc.dmap) irwls_dev_free(g_mc.dmap);
    if (g_mc.dinv) irwls_dev_free(g_mc.dinv);
    if (g_mc.dU) irwls_dev_free(g_mc.dU);
    if (g_mc.dR) irwls_dev_free(g_mc.dR);
    if (g_mc.dSinv) irwls_dev_free(g_mc.dSinv);
    free(g_mc.S1);
    free(g_mc.map);
    free(g_mc.inv);
    free(g_mc.mark);
    memset(&g_mc, 0, sizeof(g_mc));
}

static int refactor(const double *H, int n, int ldh, const int32_t *S1, long gen, irwls_gpu_stats *st) {
    const int info = irwls_dev_potrf_inv_f32(H, n, ldh, g_mc.F, 2 * n, g_mc.ZT, n);
    st->factorizations++;
    if (info != 0) {
        g_mc.valid = 0;
        return 0;
    }
    g_mc.valid = 1;
    g_mc.n = n;
    g_mc.gen = gen;
    memcpy(g_mc.S1, S1, (size_t)n * sizeof(int32_t));
    return 1;
}

/* IRWLS_MIXED_SCHUR=0 turns the exact correction for removed rows off (the
   mapped preconditioner then leaves them as rank-one differences). */
static int schur_enabled(void) {
    const char *e = getenv("IRWLS_MIXED_SCHUR");
    return !(e && *e && atoi(e) == 0);
}

/* Positions of the new free set in the factored one.  Returns the number of
   changed rows (removed + added); fills map/inv on the host. */
static int map_to_factor(const int32_t *S1, int n) {
    const int nf = g_mc.n;
    int top = 0;
    for (int f = 0; f < nf; f++)
        if (g_mc.S1[f] + 1 > top) top = g_mc.S1[f] + 1;
    for (int i = 0; i < n; i++)
        if (S1[i] + 1 > top) top = S1[i] + 1;
    if (g_mc.mark_cap < top) {
        free(g_mc.mark);
        g_mc.mark = (int32_t *)irwls_xcalloc((size_t)top, sizeof(int32_t));
        g_mc.mark_cap = top;
    }
    for (int f = 0; f < nf; f++) g_mc.mark[g_mc.S1[f]] = f + 1;
    for (int f = 0; f < nf; f++) g_mc.inv[f] = -1;
    int kept = 0, added = 0;
    for (int i = 0; i < n; i++) {
        const int f = g_mc.mark[S1[i]] - 1;
        g_mc.map[i] = f;
        if (f >= 0) {
            g_mc.inv[f] = i;
            kept++;
        } else {
            added++;
        }
    }
    for (int f = 0; f < nf; f++) g_mc.mark[g_mc.S1[f]] = 0;
    g_mc.removed = nf - kept;
    g_mc.added = added;
    if (g_mc.removed <= SCHUR_MAX) {
        int m = 0;
        for (int f = 0; f < nf; f++)
            if (g_mc.inv[f] < 0) g_mc.hR[m++] = f;
    }
    return (nf - kept) + added;
}

/* Solves with H (device, N = nS1+1, leading dimension N+1, unfactored) for
   the right-hand side whose first nS1 entries are d_rhs (device) and whose
   last entry is `last`; d_ys = y over S1 (device).  Writes the N entries of
   the solution to d_x (device) and returns 1, or returns 0 when the caller
   should use the FP64 path. */
int irwls_kkt_solve_mixed(const double *H, int nS1, const int32_t *S1, long gen, const double *d_rhs,
                          const double *d_ys, double last, double *d_x, irwls_gpu_stats *st) {
    const int n = nS1;
    const int ldh = nS1 + 2;
    if (n < 128) return 0;
    const double tol = irwls_opts()->refine_tol > 0.0 ? irwls_opts()->refine_tol : 1e-10;
    int mode = 0;   /* 0: refactor, 1: same index set, 2: mapped, 3: mapped with the removed rows corrected */
    int changes = 0;
    const int small = 4 + (g_mc.valid ? g_mc.n : n) / 256;
    if (g_mc.valid && g_mc.gen == gen && g_mc.cap >= n) {
        if (g_mc.n == n && memcmp(g_mc.S1, S1, (size_t)n * sizeof(int32_t)) == 0) {
            mode = 1;
        } else {
            changes = map_to_factor(S1, n);
            /* larger removals come early in a working set, while the weights
               (the diagonal) still change much: corrected, those solves ran
               past their CG budget and refactored anyway */
            if (schur_enabled() && g_mc.schur_off != gen && g_mc.removed > 0 && g_mc.removed <= SCHUR_MAX &&
                g_mc.added <= small)
                mode = 3;
            else if (changes <= small)
                mode = 2;
        }
    }
    if (mode == 3) {
        irwls_dev_h2d(g_mc.dR, g_mc.hR, (size_t)g_mc.removed * sizeof(int32_t));
        /* no wait for the tile's info: an unusable correction shows as a
           solve that misses its target, which refactors below */
        (void)irwls_dev_pcg_schur_setup(g_mc.ZT, g_mc.n, g_mc.n, g_mc.dR, g_mc.removed, g_mc.dU, g_mc.dSinv, 0);
    }
    if (mode == 0) {
        mc_reserve(n);
        if (!refactor(H, n, ldh, S1, gen, st)) return 0;
    }
    /* right-hand sides: r1 and the border y */
    irwls_dev_d2d(g_mc.B, d_rhs, (size_t)n * sizeof(double));
    irwls_dev_d2d(g_mc.B + n, d_ys, (size_t)n * sizeof(double));
    double rel[2];
    int it;
    const int nf = g_mc.n;
    if (mode == 2 || mode == 3) {
        irwls_dev_h2d(g_mc.dmap, g_mc.map, (size_t)n * sizeof(int32_t));
        irwls_dev_h2d(g_mc.dinv, g_mc.inv, (size_t)nf * sizeof(int32_t));
        if (mode == 3)
            it = irwls_dev_pcg_spd2_schur(H, n, ldh, g_mc.B, g_mc.F, 2 * nf, g_mc.ZT, nf, nf, g_mc.dmap, g_mc.dinv,
                                          g_mc.dR, g_mc.removed, g_mc.dU, g_mc.dSinv, tol, 16 + 2 * g_mc.added,
                                          g_mc.X, rel);
        else
            it = irwls_dev_pcg_spd2_mapped(H, n, ldh, g_mc.B, g_mc.F, 2 * nf, g_mc.ZT, nf, nf, g_mc.dmap, g_mc.dinv,
                                           tol, 16 + 2 * changes, g_mc.X, rel);
    } else {
        it = irwls_dev_pcg_spd2(H, n, ldh, g_mc.B, g_mc.F, 2 * n, g_mc.ZT, n, tol, mode ? 16 : 200, g_mc.X, rel);
    }
    st->pcg_iters += it < 0 ? -it : it;
    /* where the diagonal's change dominates, the correction saves a
       factorization but costs more CG iterations than it was worth */
    if (mode == 3 && (it < 0 || it > SCHUR_ITS)) g_mc.schur_off = gen;
    if (it < 0 && mode != 0) {
        /* the earlier factor no longer preconditions well enough */
        mc_reserve(n);
        if (!refactor(H, n, ldh, S1, gen, st)) return 0;
        it = irwls_dev_pcg_spd2(H, n, ldh, g_mc.B, g_mc.F, 2 * n, g_mc.ZT, n, tol, 200, g_mc.X, rel);
        st->pcg_iters += it < 0 ? -it : it;
    }
    if (it < 0) {
        g_mc.valid = 0;
        return 0;
    }
    /* the closed form on the device, into the caller's solution vector */
    irwls_dev_kkt_closed_form(n, g_mc.X, d_ys, last, d_x);
    return 1;
}
