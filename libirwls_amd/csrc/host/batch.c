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
 * waits.  A round of a size class (systems up to 512 rows are batched apart
 * from larger ones: a round lasts as long as its largest system) starts
 * when every problem of that class that is inside an inner loop is queued
 * or already in a running round, or when the earliest queued one has waited
 * a set time (IRWLS_BATCH_WAIT_US, default 2500 microseconds); problems
 * busy with an outer iteration (kernel products, working-set selection) are
 * not waited for.  The round runs the queued iterations together
 * (irwls_dev_inner_batched_slot: one replay of a captured launch sequence
 * that reads each problem's operands from a device table, tile shape chosen
 * by its autotuner), on the stream of the thread that starts it.  Up to
 * BATCH_SLOTS rounds run at once, so a problem that queues while a round
 * runs need not wait for it.  An iteration that no other problem would join
 * (none of its class queued or between two iterations) runs the
 * single-problem kernels at once instead (IRWLS_BATCH_ALONE=0: never).
 * Each problem waits on a condition variable of its own: a round's end
 * wakes its own problems, and only the earliest queued problem of a class
 * waits for the deadline and is woken when a round may start.  The
 * per-problem arithmetic is that of the standalone solve, so each problem's
 * result is bit-identical to its standalone trainFULL run.
 *
 * IRWLS_BATCH_TRACE=<file> appends a line per round (host clock, problems,
 * class, largest system, round slot, queue wait) and per problem (finish
 * time, iterations, and its time in the stream sync, the queue, rounds and
 * the wake-up after them).
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
    int nprob, active, npend, rounds;
    int alone_ok, alone;  /* iterations run alone when no other would join (IRWLS_BATCH_ALONE=0: never) */
    int inflight;         /* problems in running batches (lock released meanwhile) */
    unsigned busy;        /* round slots in use */
    double wait_s;        /* longest wait before a partial batch runs */
    int *pending;         /* 0 idle, 1 queued, 2 in a running batch */
    pthread_cond_t *cvp;  /* per problem: its batch has ended, or (the earliest queued one of a
                             size class) a batch of its class may now start */
    pthread_cond_t cv_exit;   /* workers that have finished wait here for the last one */
    int *inner;           /* 1 while a problem is inside an inner IRWLS loop */
    int *cls;             /* size class of each problem's last queued iteration */
    double *t_queue;      /* when each queued problem queued */
    irwls_inner_work *work;   /* each queued problem's inner iteration */
    /* IRWLS_BATCH_TRACE=<file>: one line per round (start and end on the
       host clock relative to the call, problems, size class, largest
       system, round slot, summed queue wait of its problems) */
    double t0;
    double *t_done;          /* when each problem finished (with the trace only) */
    double *t_in;            /* time each problem spent queued or in a batch */
    double *t_part;          /* per problem: stream sync, queue wait, round, wake-up (4 sums) */
    double *t_rend;          /* end of the round that ran each problem's last iteration */
    int *n_alone;            /* iterations each problem ran alone */
    double *trace;
    int ntrace, trace_cap;
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

/* The earliest queued problem of class c (it waits for the batching
   deadline and is woken when a batch may start), or -1. */
static int leader_of(const batch_sched *s, int c) {
    int b = -1;
    for (int i = 0; i < s->nprob; i++)
        if (s->pending[i] == 1 && s->cls[i] == c && (b < 0 || s->t_queue[i] < s->t_queue[b])) b = i;
    return b;
}

static void wake_leaders(batch_sched *s) {
    for (int c = 0; c < 2; c++) {
        const int b = leader_of(s, c);
        if (b >= 0) pthread_cond_signal(&s->cvp[b]);
    }
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
    const double tb = irwls_now();
    double qwait = 0.0;
    int nmax = 0;
    for (int j = 0; j < m; j++) {
        qwait += tb - s->t_queue[ids[j]];
        if (w[j].N > nmax) nmax = w[j].N;
    }
    pthread_mutex_unlock(&s->mu);
    irwls_dev_inner_batched_slot(slot, m, w);
    irwls_dev_sync();
    const double te = irwls_now();
    pthread_mutex_lock(&s->mu);
    if (s->t_part)
        for (int j = 0; j < m; j++) {
            s->t_part[4 * ids[j] + 1] += tb - s->t_queue[ids[j]];
            s->t_part[4 * ids[j] + 2] += te - tb;
            s->t_rend[ids[j]] = te;
        }
    if (s->trace) {
        if (s->ntrace == s->trace_cap) {
            s->trace_cap *= 2;
            s->trace = (double *)realloc(s->trace, (size_t)s->trace_cap * 7 * sizeof(double));
            if (!s->trace) {
                fprintf(stderr, "libirwls: trainFULLBatched: out of host memory for the round trace\n");
                exit(2);
            }
        }
        double *r = s->trace + (size_t)s->ntrace++ * 7;
        r[0] = tb - s->t0;
        r[1] = te - s->t0;
        r[2] = m;
        r[3] = c;
        r[4] = nmax;
        r[5] = slot;
        r[6] = qwait;
    }
    for (int j = 0; j < m; j++) {
        s->pending[ids[j]] = 0;
        pthread_cond_signal(&s->cvp[ids[j]]);   /* only this batch's problems wake */
    }
    s->inflight -= m;
    s->busy &= ~(1u << slot);
    s->rounds++;
    free(ids);
    free(w);
    wake_leaders(s);   /* a round slot is free again */
}

/* Problems of class c inside an inner loop that are neither queued nor in a
   running batch: they are between two iterations and will queue shortly.
   Problems outside an inner loop (an outer iteration's kernel products and
   working-set selection, milliseconds for the larger ones) are not waited
   for. */
static int soon_class(const batch_sched *s, int c) {
    int k = 0;
    for (int i = 0; i < s->nprob; i++)
        if (s->inner[i] && s->pending[i] == 0 && s->cls[i] == c) k++;
    return k;
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
    pthread_mutex_lock(&s->mu);
    s->inner[me] = 1;
    s->cls[me] = size_class_of(w->N);
    /* soon_class counts this problem too (not yet queued) */
    if (s->alone_ok && npend_class(s, s->cls[me]) == 0 && soon_class(s, s->cls[me]) == 1) {
        /* no other problem would join: the single-problem kernels, at once
           (the same arithmetic, so the same values) */
        s->alone++;
        if (s->n_alone) s->n_alone[me]++;
        pthread_mutex_unlock(&s->mu);
        return 0;
    }
    pthread_mutex_unlock(&s->mu);
    const double t_enter = irwls_now();
    irwls_dev_sync();   /* the iteration's inputs are complete on this thread's stream */
    const double t_synced = irwls_now();
    pthread_mutex_lock(&s->mu);
    if (s->t_part) s->t_part[4 * me] += t_synced - t_enter;
    s->work[me] = *w;
    s->pending[me] = 1;
    s->t_queue[me] = irwls_now();
    s->npend++;
    while (s->pending[me] != 0) {
        /* a thread whose own problem is queued starts the next batch, on a
           free round slot; while it cannot, the earliest queued problem of
           the class waits for the deadline and the others for their batch */
        const int slot = (s->pending[me] == 1) ? free_slot(s) : -1;
        const int c = s->cls[me];
        if (slot >= 0) {
            const double due = first_queued(s, c) + s->wait_s;
            if (soon_class(s, c) == 0 || irwls_now() >= due) {
                run_batch(s, slot, c);
                continue;
            }
        }
        if (slot >= 0 && leader_of(s, c) == me) {
            const double due = s->t_queue[me] + s->wait_s;
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
            pthread_cond_timedwait(&s->cvp[me], &s->mu, &ts);
        } else {
            pthread_cond_wait(&s->cvp[me], &s->mu);
        }
    }
    if (s->t_in) {
        const double now = irwls_now();
        s->t_in[me] += now - t_enter;
        s->t_part[4 * me + 3] += now - s->t_rend[me];
    }
    pthread_mutex_unlock(&s->mu);
    return 1;
}

void irwls_batch_inner_end(void) {
    batch_sched *s = t_sched;
    if (!s) return;
    pthread_mutex_lock(&s->mu);
    s->inner[t_slot] = 0;
    wake_leaders(s);   /* a waiting batch may no longer need to wait for this one */
    pthread_mutex_unlock(&s->mu);
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
    if (s->t_done) s->t_done[j->slot] = irwls_now() - s->t0;
    t_sched = NULL;
    t_rng = NULL;
    irwls_mixed_release();
    irwls_dev_thread_end();
    pthread_mutex_lock(&s->mu);
    s->active--;
    s->inner[j->slot] = 0;
    wake_leaders(s);   /* the waiting ones may now be all that is left */
    if (s->active == 0) pthread_cond_broadcast(&s->cv_exit);
    /* The thread stays until every problem has finished: nothing a worker's
       HIP calls created or registered (on its stack, in its thread-local
       storage, or in a profiler's per-thread state) goes away while other
       workers still launch. */
    while (s->active > 0) pthread_cond_wait(&s->cv_exit, &s->mu);
    pthread_mutex_unlock(&s->mu);
    return NULL;
}

int trainFULLBatched(const svm_dataset *datasets, const properties *props, int nprob, double **out) {
    if (nprob <= 0) return 0;
    const double t0 = irwls_now();
    irwls_dev_init();
    batch_sched s;
    memset(&s, 0, sizeof(s));
    s.t0 = t0;
    const char *trace_path = getenv("IRWLS_BATCH_TRACE");
    if (trace_path && *trace_path) {
        s.trace_cap = 1024;
        s.trace = (double *)irwls_xmalloc((size_t)s.trace_cap * 7 * sizeof(double));
        s.t_done = (double *)irwls_xcalloc((size_t)nprob, sizeof(double));
        s.t_in = (double *)irwls_xcalloc((size_t)nprob, sizeof(double));
        s.n_alone = (int *)irwls_xcalloc((size_t)nprob, sizeof(int));
        s.t_part = (double *)irwls_xcalloc((size_t)nprob * 4, sizeof(double));
        s.t_rend = (double *)irwls_xcalloc((size_t)nprob, sizeof(double));
    }
    pthread_mutex_init(&s.mu, NULL);
    pthread_cond_init(&s.cv_exit, NULL);
    s.cvp = (pthread_cond_t *)irwls_xcalloc((size_t)nprob, sizeof(pthread_cond_t));
    for (int p = 0; p < nprob; p++) pthread_cond_init(&s.cvp[p], NULL);
    s.nprob = nprob;
    s.active = nprob;
    s.pending = (int *)irwls_xcalloc((size_t)nprob, sizeof(int));
    s.t_queue = (double *)irwls_xcalloc((size_t)nprob, sizeof(double));
    s.inner = (int *)irwls_xcalloc((size_t)nprob, sizeof(int));
    s.cls = (int *)irwls_xcalloc((size_t)nprob, sizeof(int));
    {
        const char *w = getenv("IRWLS_BATCH_WAIT_US");
        s.wait_s = (w && *w) ? atof(w) * 1e-6 : 2.5e-3;
    }
    s.work = (irwls_inner_work *)irwls_xcalloc((size_t)nprob, sizeof(irwls_inner_work));
    {
        const char *a = getenv("IRWLS_BATCH_ALONE");
        s.alone_ok = !(a && *a && atoi(a) == 0);
    }
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
    if (s.trace) {
        FILE *f = fopen(trace_path, "a");
        if (f) {
            fprintf(f, "# call: %d problems, %.6f s, %d iterations alone; columns: start end problems class largest_n slot "
                       "queue_wait_sum\n",
                    nprob, tot.t_total, s.alone);
            for (int i = 0; i < s.ntrace; i++) {
                const double *r = s.trace + (size_t)i * 7;
                fprintf(f, "%.6f %.6f %d %d %d %d %.6f\n", r[0], r[1], (int)r[2], (int)r[3], (int)r[4], (int)r[5], r[6]);
            }
            for (int p = 0; p < nprob; p++)
                fprintf(f, "# problem %d: done at %.6f s, %d inner and %d outer iterations, %d alone, %.6f s queued "
                           "or in batches (sync %.6f, queue %.6f, round %.6f, wake %.6f)\n",
                        p, s.t_done[p], jobs[p].stats.inner_iters, jobs[p].stats.outer_iters, s.n_alone[p], s.t_in[p],
                        s.t_part[4 * p], s.t_part[4 * p + 1], s.t_part[4 * p + 2], s.t_part[4 * p + 3]);
            fclose(f);
        } else {
            fprintf(stderr, "libirwls: trainFULLBatched: cannot write the round trace %s\n", trace_path);
        }
        free(s.trace);
        free(s.t_done);
        free(s.t_in);
        free(s.t_part);
        free(s.t_rend);
        free(s.n_alone);
    }
    for (int p = 0; p < nprob; p++) pthread_cond_destroy(&s.cvp[p]);
    free(s.cvp);
    pthread_cond_destroy(&s.cv_exit);
    pthread_mutex_destroy(&s.mu);
    free(s.pending);
    free(s.t_queue);
    free(s.inner);
    free(s.cls);
    free(s.work);
    free(jobs);
    free(th);
    return 0;
}
