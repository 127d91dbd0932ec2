/*
 * Host-side parallel loops: contiguous ranges of [0, n) on up to 16
 * threads (the calling thread takes the first range).  Used where the host
 * walks every row of a dataset (device upload, range scans).
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "internal.h"

typedef struct {
    irwls_range_fn fn;
    void *arg;
    int t;
    long begin, end;
} range_job;

static void *run_range(void *p) {
    range_job *j = (range_job *)p;
    j->fn(j->arg, j->t, j->begin, j->end);
    return NULL;
}

int irwls_host_threads(long work, long grain) {
    if (irwls_batch_worker()) return 1;   /* a batched run already has a thread per problem */
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    long nt = grain > 0 ? work / grain : 1;
    if (nt > 16) nt = 16;
    if (nt > ncpu) nt = ncpu;
    if (nt < 1) nt = 1;
    return (int)nt;
}

void irwls_parallel_ranges(long n, int nt, irwls_range_fn fn, void *arg) {
    if (n <= 0) return;
    if (nt > n) nt = (int)n;
    if (nt <= 1) {
        fn(arg, 0, 0, n);
        return;
    }
    range_job jobs[16];
    pthread_t th[16];
    if (nt > 16) nt = 16;
    for (int t = 0; t < nt; t++) {
        jobs[t].fn = fn;
        jobs[t].arg = arg;
        jobs[t].t = t;
        jobs[t].begin = n * t / nt;
        jobs[t].end = n * (t + 1) / nt;
    }
    for (int t = 1; t < nt; t++)
        if (pthread_create(&th[t], NULL, run_range, &jobs[t]) != 0) {
            fprintf(stderr, "libirwls: cannot start a host worker thread\n");
            exit(2);
        }
    run_range(&jobs[0]);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
}
