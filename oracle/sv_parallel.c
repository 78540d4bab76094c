/*
 * sv_parallel.c -- the reference's multi-threaded CPU path, restated for the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it; nothing in the product links or calls it.
 *
 * StatevectorParallelExecutor.forRange (engines/.../statevector/StatevectorParallelExecutor.java:29-73) splits
 * [0, work) into leaves of max(minimumWorkUnits, ceil(work / (8 P))) and runs them on a ForkJoin pool of P
 * workers; here P POSIX threads take the same leaves from a shared counter.  The per-pair arithmetic is the
 * scalar form of the reference's pair kernels: applySingleGatePairs / applySingleGatePair
 * (StatevectorOps.java:749-784, :830-849) over pair indices expanded with a zero bit at the target
 * (expandIndexWithZeroBit :984-989), and the parallel applyCnot (:147-172) over the same pairs, exchanging
 * where the control bit is set.  Compiled with -ffp-contract=off like sv_oracle.c.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <string.h>

#include "../include/statecraft_hip.h"

typedef struct {
    double *state;
    int kind; /* 0: single-qubit gate, 1: CNOT */
    int target, control;
    double g[8];
    uint64_t work, leaf;
    atomic_uint_fast64_t next;
} par_job;

static inline uint64_t expand_zero_bit(uint64_t k, int bit) {
    const uint64_t low = k & ((1ULL << bit) - 1ULL);
    return ((k >> bit) << (bit + 1)) | low;
}

static void run_range(par_job *j, uint64_t lo, uint64_t hi) {
    double *s = j->state;
    const uint64_t stride = 1ULL << j->target;
    if (j->kind == 0) {
        const double g00r = j->g[0], g00i = j->g[1], g01r = j->g[2], g01i = j->g[3];
        const double g10r = j->g[4], g10i = j->g[5], g11r = j->g[6], g11i = j->g[7];
        for (uint64_t k = lo; k < hi; k++) {
            const uint64_t i0 = expand_zero_bit(k, j->target), i1 = i0 | stride;
            const double a0r = s[2 * i0], a0i = s[2 * i0 + 1], a1r = s[2 * i1], a1i = s[2 * i1 + 1];
            s[2 * i0] = g00r * a0r - g00i * a0i + g01r * a1r - g01i * a1i;
            s[2 * i0 + 1] = g00r * a0i + g00i * a0r + g01r * a1i + g01i * a1r;
            s[2 * i1] = g10r * a0r - g10i * a0i + g11r * a1r - g11i * a1i;
            s[2 * i1 + 1] = g10r * a0i + g10i * a0r + g11r * a1i + g11i * a1r;
        }
    } else {
        const uint64_t cm = 1ULL << j->control;
        for (uint64_t k = lo; k < hi; k++) {
            const uint64_t i0 = expand_zero_bit(k, j->target);
            if (!(i0 & cm)) continue;
            const uint64_t i1 = i0 | stride;
            double tr = s[2 * i0], ti = s[2 * i0 + 1];
            s[2 * i0] = s[2 * i1];
            s[2 * i0 + 1] = s[2 * i1 + 1];
            s[2 * i1] = tr;
            s[2 * i1 + 1] = ti;
        }
    }
}

/* A persistent pool (ForkJoinPool.commonPool-like: created once, reused by every op).  The calling thread
 * works too; a job ends when every leaf is done. */
static pthread_mutex_t pool_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t pool_cv = PTHREAD_COND_INITIALIZER, done_cv = PTHREAD_COND_INITIALIZER;
static par_job *pool_job = NULL;
static uint64_t pool_gen = 0;
static int pool_size = 0, pool_busy = 0;

static void drain(par_job *j) {
    for (;;) {
        const uint64_t lo = atomic_fetch_add(&j->next, j->leaf);
        if (lo >= j->work) break;
        run_range(j, lo, lo + j->leaf < j->work ? lo + j->leaf : j->work);
    }
}

static void *pool_worker(void *arg) {
    (void)arg;
    uint64_t seen = 0;
    pthread_mutex_lock(&pool_mu);
    for (;;) {
        while (pool_gen == seen) pthread_cond_wait(&pool_cv, &pool_mu);
        seen = pool_gen;
        par_job *j = pool_job;
        if (!j) continue;  /* woke after that job had already ended */
        pool_busy++;
        pthread_mutex_unlock(&pool_mu);
        drain(j);
        pthread_mutex_lock(&pool_mu);
        if (--pool_busy == 0) pthread_cond_broadcast(&done_cv);
    }
    return NULL;
}

static void for_range(par_job *j, int threads, uint64_t min_work) {
    const uint64_t tasks = (uint64_t)(threads > 0 ? threads : 1) * 8;
    const uint64_t balanced = (j->work + tasks - 1) / tasks;
    j->leaf = balanced > min_work ? balanced : min_work;
    atomic_store(&j->next, 0);
    if (threads <= 1 || j->work < min_work) {
        run_range(j, 0, j->work);
        return;
    }
    pthread_mutex_lock(&pool_mu);
    while (pool_size < threads - 1 && pool_size < 255) {
        pthread_t t;
        if (pthread_create(&t, NULL, pool_worker, NULL) != 0) break;
        pthread_detach(t);
        pool_size++;
    }
    pool_job = j;
    pool_gen++;
    pthread_cond_broadcast(&pool_cv);
    pthread_mutex_unlock(&pool_mu);
    drain(j);
    pthread_mutex_lock(&pool_mu);
    /* workers that woke for this generation finish before the job (on the caller's stack) goes away */
    while (pool_busy > 0) pthread_cond_wait(&done_cv, &pool_mu);
    pool_job = NULL;
    pthread_mutex_unlock(&pool_mu);
}

/* The single-qubit and CNOT ops of a lowered circuit on `threads` threads (other kinds: ORC_E_ARG = 2). */
int orc_run_ops_parallel(double *state, int n, const sc_op *ops, int64_t n_ops, const double *mats, int threads,
                         int64_t min_work) {
    for (int64_t i = 0; i < n_ops; i++) {
        par_job j;
        memset(&j, 0, sizeof j);
        j.state = state;
        j.work = 1ULL << (n - 1);
        if (ops[i].kind == SC_OP_SINGLE) {
            j.kind = 0;
            j.target = ops[i].q0;
            memcpy(j.g, mats + 2 * (int64_t)ops[i].mat, sizeof j.g);
        } else if (ops[i].kind == SC_OP_CNOT) {
            j.kind = 1;
            j.control = ops[i].q0;
            j.target = ops[i].q1;
        } else {
            return 2;
        }
        for_range(&j, threads, (uint64_t)min_work);
    }
    return 0;
}

/* One single-qubit gate (g: 4 complex, row-major) on `threads` threads. */
void orc_apply_single_parallel(double *state, int n, int target, const double *g, int threads, int64_t min_work) {
    par_job j;
    memset(&j, 0, sizeof j);
    j.state = state;
    j.work = 1ULL << (n - 1);
    j.target = target;
    memcpy(j.g, g, sizeof j.g);
    for_range(&j, threads, (uint64_t)min_work);
}

/* StatevectorKernelMicrobenchmark's timed loop: `iterations` passes over `count` single-qubit gates (targets,
 * 4 complex each) on `threads` threads (1: the serial runAoS). */
void orc_gate_loop(double *state, int n, const int32_t *targets, const double *gates, int count, int iterations,
                   int threads, int64_t min_work) {
    for (int it = 0; it < iterations; it++)
        for (int k = 0; k < count; k++) orc_apply_single_parallel(state, n, targets[k], gates + 8 * k, threads, min_work);
}

