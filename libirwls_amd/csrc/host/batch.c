/*
 * Batched full training: many independent trainFULL problems (one-vs-rest
 * tasks, C/gamma grid points) in one call, their inner IRWLS iterations
 * batched.
 *
 * Each problem runs the unchanged trainFULL control flow on a host thread of
 * its own, with a device stream, scratch and staging buffers of its own
 * (irwls_dev_thread_begin), so the problems' kernel-block products overlap
 * on the device.  The device work of each inner IRWLS iteration
 * (src/full-train.c:158-249: right-hand side, KKT assembly, the solve, the
 * weights after it and the error update) runs in batched rounds: a problem
 * that reaches it finishes its stream's work, queues the iteration and
 * waits; when every problem that is not already in a running round is
 * waiting, or when the earliest waiting one has waited a set time
 * (IRWLS_BATCH_WAIT_US, default 1000 microseconds), one round runs the
 * waiting ones' iterations together (irwls_dev_inner_batched_slot: one
 * replay of a captured launch sequence that reads each problem's operands
 * from a device table, tile shape chosen by its autotuner), and they
 * continue.  Systems up to 512 rows are batched apart from larger ones.  Up
 * to
 * BATCH_SLOTS rounds run at once, each on the stream of the waiting problem
 * whose thread started it, so a problem that queues while a batch runs need
 * not wait for that batch.  The time limit keeps a problem busy with a longer
 * step (an outer iteration's kernel products) from holding back the others.
 * The per-problem arithmetic is that of the standalone solve, so each
 * problem's result is bit-identical to its standalone trainFULL run.
 *
 * Each problem draws its working-set permutations (src/full-train.c:631)
 * from a private stream with glibc's rand() algorithm seeded like srand(0)
 * (the reference's full-train main calls srand(0) before trainFULL), so it
 * does not depend on the other problems' progress.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "internal.h"
#include "full-train.h"

#define BATCH_SLOTS 4     /* batches in flight at once (round slots of the device solver) */
#define BATCH_SMALL_N 512 /* systems up to this size are batched apart from the larger ones */

typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    int nprob, active, npend, rounds;
    int inflight;         /* problems in running batches (lock released meanwhile) */
    unsigned busy;        /* round slots in use */
    double wait_s;        /* longest wait before a partial batch runs */
    int *pending;         /* 0 idle, 1 queued, 2 in a running batch */
    double *t_queue;      /* when each queued problem queued */
    irwls_inner_work *work;   /* each queued problem's inner iteration */
} batch_sched;

typedef struct {
    batch_sched *s;
    int slot;
    svm_dataset ds;
    properties props;
    double *out;
    irwls_gpu_stats stats;
    struct random_data rd;
    char rstate[128];
} batch_job;

static __thread batch_sched *t_sched;
static __thread int t_slot;
static __thread struct random_data *t_rng;

int irwls_batch_worker(void) { return t_sched != NULL; }

/* A block of libc rand() draws without its lock (about 25 ns a call): libc
   is switched to a scratch state while the caller's state is stepped with
   random_r, the reentrant form of the same generator (rand() is random()),
   and the state is handed back with its rear position advanced, so the
   draws and every later rand() are exactly those of calling rand() as many
   times.  The rear position travels in the state's header word, as
   initstate/setstate store it.  In a batched run the worker's private
   stream is used directly. */
typedef struct {
    struct random_data rd;
    char own[128], scratch[128];
    char *caller;
    int active;
} rand_block;
_Static_assert(sizeof(rand_block) <= sizeof(irwls_rand_block), "irwls_rand_block too small");

void irwls_rand_begin(irwls_rand_block *ob) {
    rand_block *b = (rand_block *)ob;
    memset(b, 0, sizeof(*b));
    if (t_rng) return;
    initstate_r(1u, b->own, sizeof(b->own), &b->rd);              /* a valid buffer for setstate_r to leave */
    b->caller = initstate(1u, b->scratch, sizeof(b->scratch));    /* libc on scratch; caller's position saved */
    setstate_r(b->caller, &b->rd);                                /* step the caller's state here */
    b->active = 1;
}

int irwls_rand_draw(irwls_rand_block *ob) {
    rand_block *b = (rand_block *)ob;
    int32_t r = 0;
    if (b->active) random_r(&b->rd, &r);
    else if (t_rng) random_r(t_rng, &r);
    else r = rand();
    return r;
}

void irwls_rand_end(irwls_rand_block *ob) {
    rand_block *b = (rand_block *)ob;
    if (!b->active) return;
    /* the header word: rear position and generator type, as setstate reads it */
    int32_t *st = b->rd.state;
    if (b->rd.rand_type != 0) st[-1] = 5 * (int32_t)(b->rd.rptr - st) + b->rd.rand_type;   /* 5 = MAX_TYPES */
    setstate(b->caller);
    b->active = 0;
}

int irwls_rand(void) {
    if (t_rng) {
        int32_t r = 0;
        random_r(t_rng, &r);
        return r;
    }
    return rand();
}

/* A free round slot, or -1. */
static int free_slot(const batch_sched *s) {
    for (int k = 0; k < BATCH_SLOTS; k++)
        if (!((s->busy >> k) & 1u)) return k;
    return -1;
}

/* Size class of a system: a batch runs as long as its largest system, so
   small systems are batched among themselves. */
static int size_class_of(int n) { return n > BATCH_SMALL_N ? 1 : 0; }

/* Queued problems of class c. */
static int npend_class(const batch_sched *s, int c) {
    int k = 0;
    for (int i = 0; i < s->nprob; i++)
        if (s->pending[i] == 1 && size_class_of(s->work[i].N) == c) k++;
    return k;
}

/* Runs the queued solves of size class c on round slot `slot`.  Called with
   the lock held; releases it during the device work and returns with it
   held. */
static void run_batch(batch_sched *s, int slot, int c) {
    const int k = npend_class(s, c);
    int *ids = (int *)irwls_xmalloc((size_t)k * sizeof(int));
    irwls_inner_work *w = (irwls_inner_work *)irwls_xmalloc((size_t)k * sizeof(irwls_inner_work));
    int m = 0;
    for (int i = 0; i < s->nprob && m < k; i++) {
        if (s->pending[i] != 1 || size_class_of(s->work[i].N) != c) continue;
        ids[m] = i;
        w[m] = s->work[i];
        s->pending[i] = 2;
        m++;
    }
    s->npend -= m;
    s->inflight += m;
    s->busy |= 1u << slot;
    pthread_mutex_unlock(&s->mu);
    irwls_dev_inner_batched_slot(slot, m, w);
    irwls_dev_sync();
    pthread_mutex_lock(&s->mu);
    for (int j = 0; j < m; j++) s->pending[ids[j]] = 0;
    s->inflight -= m;
    s->busy &= ~(1u << slot);
    s->rounds++;
    free(ids);
    free(w);
    pthread_cond_broadcast(&s->cv);
}

/* Earliest queue time of the queued problems of class c. */
static double first_queued(const batch_sched *s, int c) {
    double t = 1e300;
    for (int i = 0; i < s->nprob; i++)
        if (s->pending[i] == 1 && size_class_of(s->work[i].N) == c && s->t_queue[i] < t) t = s->t_queue[i];
    return t;
}

int irwls_batch_inner(const irwls_inner_work *w) {
    batch_sched *s = t_sched;
    if (!s) return 0;
    const int me = t_slot;
    irwls_dev_sync();   /* the iteration's inputs are complete on this thread's stream */
    pthread_mutex_lock(&s->mu);
    s->work[me] = *w;
    s->pending[me] = 1;
    s->t_queue[me] = irwls_now();
    s->npend++;
    while (s->pending[me] != 0) {
        /* a thread whose own problem is queued starts the next batch, on a free round slot */
        const int slot = (s->pending[me] == 1) ? free_slot(s) : -1;
        if (slot >= 0) {
            const int c = size_class_of(s->work[me].N);
            const double due = first_queued(s, c) + s->wait_s;
            if (s->npend == s->active - s->inflight || irwls_now() >= due) {
                run_batch(s, slot, c);
                continue;
            }
            struct timespec ts;
            clock_gettime(CLOCK_REALTIME, &ts);
            const double left = due - irwls_now();
            const long ns = (long)((left > 0.0 ? left : 0.0) * 1e9) + 1000;
            ts.tv_sec += ns / 1000000000L;
            ts.tv_nsec += ns % 1000000000L;
            if (ts.tv_nsec >= 1000000000L) {
                ts.tv_sec++;
                ts.tv_nsec -= 1000000000L;
            }
            pthread_cond_timedwait(&s->cv, &s->mu, &ts);
        } else {
            pthread_cond_wait(&s->cv, &s->mu);
        }
    }
    pthread_mutex_unlock(&s->mu);
    return 1;
}

static void *worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    batch_sched *s = j->s;
    irwls_dev_thread_begin();
    t_sched = s;
    t_slot = j->slot;
    t_rng = &j->rd;
    (void)irwls_dev_wait_seconds(1);
    j->out = trainFULL(j->ds, j->props);
    irwls_gpu_last_stats(&j->stats);
    j->stats.t_wait = irwls_dev_wait_seconds(1);
    t_sched = NULL;
    t_rng = NULL;
    irwls_mixed_release();
    irwls_dev_thread_end();
    pthread_mutex_lock(&s->mu);
    s->active--;
    pthread_cond_broadcast(&s->cv);   /* the waiting ones may now be all that is left */
    /* The thread stays until every problem has finished: nothing a worker's
       HIP calls created or registered (on its stack, in its thread-local
       storage, or in a profiler's per-thread state) goes away while other
       workers still launch. */
    while (s->active > 0) pthread_cond_wait(&s->cv, &s->mu);
    pthread_mutex_unlock(&s->mu);
    return NULL;
}

int trainFULLBatched(const svm_dataset *datasets, const properties *props, int nprob, double **out) {
    if (nprob <= 0) return 0;
    const double t0 = irwls_now();
    irwls_dev_init();
    batch_sched s;
    memset(&s, 0, sizeof(s));
    pthread_mutex_init(&s.mu, NULL);
    pthread_cond_init(&s.cv, NULL);
    s.nprob = nprob;
    s.active = nprob;
    s.pending = (int *)irwls_xcalloc((size_t)nprob, sizeof(int));
    s.t_queue = (double *)irwls_xcalloc((size_t)nprob, sizeof(double));
    {
        const char *w = getenv("IRWLS_BATCH_WAIT_US");
        s.wait_s = (w && *w) ? atof(w) * 1e-6 : 1e-3;
    }
    s.work = (irwls_inner_work *)irwls_xcalloc((size_t)nprob, sizeof(irwls_inner_work));
    /* every worker's stream comes from the pool, created here on the calling thread */
    irwls_dev_reserve_streams(nprob);
    batch_job *jobs = (batch_job *)irwls_xcalloc((size_t)nprob, sizeof(batch_job));
    pthread_t *th = (pthread_t *)irwls_xcalloc((size_t)nprob, sizeof(pthread_t));
    for (int p = 0; p < nprob; p++) {
        jobs[p].s = &s;
        jobs[p].slot = p;
        jobs[p].ds = datasets[p];
        jobs[p].props = props[p];
        memset(&jobs[p].rd, 0, sizeof(jobs[p].rd));
        initstate_r(0u, jobs[p].rstate, sizeof(jobs[p].rstate), &jobs[p].rd);
    }
    for (int p = 0; p < nprob; p++) {
        if (pthread_create(&th[p], NULL, worker, &jobs[p]) != 0) {
            fprintf(stderr, "libirwls: trainFULLBatched: cannot start a worker thread\n");
            exit(2);
        }
    }
    for (int p = 0; p < nprob; p++) pthread_join(th[p], NULL);
    irwls_gpu_stats tot;
    memset(&tot, 0, sizeof(tot));
    for (int p = 0; p < nprob; p++) {
        out[p] = jobs[p].out;
        tot.outer_iters += jobs[p].stats.outer_iters;
        tot.inner_iters += jobs[p].stats.inner_iters;
        tot.solves += jobs[p].stats.solves;
        tot.factorizations += jobs[p].stats.factorizations;
        tot.pcg_iters += jobs[p].stats.pcg_iters;
        tot.fp64_fallbacks += jobs[p].stats.fp64_fallbacks;
        tot.t_solve += jobs[p].stats.t_solve;
        tot.t_wait += jobs[p].stats.t_wait;
    }
    tot.batch_rounds = s.rounds;
    tot.t_total = irwls_now() - t0;
    irwls_stats_set(&tot);
    pthread_cond_destroy(&s.cv);
    pthread_mutex_destroy(&s.mu);
    free(s.pending);
    free(s.t_queue);
    free(s.work);
    free(jobs);
    free(th);
    return 0;
}
