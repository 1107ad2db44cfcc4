/*
 * mbg_sched.h -- a worker pool that runs the mapper's per-read host code as coroutines and turns its
 * blocking per-call work into device batches.
 *
 * The reference maps one read at a time: mb_map_sai (map-algo.c:571-700) and mb_pair (pe.c:448-582)
 * call the DP and local-SW kernels one job at a time, each call depending on the previous one's
 * result (align.c:548-757: the left extension, the gap fills with their z-drop break and second pass,
 * the right extension; pe.c:319-378: the rescue candidates).  Here every read (or read pair) runs in a
 * coroutine on one of n host threads.  A hook that needs the device posts its request and yields; when
 * every live coroutine of every thread has yielded or finished, one thread runs the `round` callback,
 * which executes all posted requests as one batch, and every suspended coroutine resumes with its
 * result.  A read's own calls therefore keep their order and their data dependences exactly, while
 * the device sees the k-th pending call of every read at once.
 *
 * x86-64 System V only: the context switch saves the callee-saved registers and the stack pointer.
 */
#ifndef MBG_SCHED_H
#define MBG_SCHED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbs_pool_s mbs_pool_t;

/* Runs task i; may call mbs_yield() any number of times. tid is the pool thread (0..n_threads-1). */
typedef void (*mbs_task_f)(void *data, int64_t i, int tid);
/* Runs once per round on one pool thread while all others wait; n_threads lists may hold requests. */
typedef void (*mbs_round_f)(void *data, int n_threads);

mbs_pool_t *mbs_pool_init(int n_threads, int64_t stack_bytes);
void mbs_pool_destroy(mbs_pool_t *p);
int mbs_pool_size(const mbs_pool_t *p);

/* Runs task(data, i, tid) for every i in [0, n) and returns when all have finished.  Tasks are taken
 * in chunks of `chunk` consecutive indices.  round may be null when no task yields. */
void mbs_run(mbs_pool_t *p, int64_t n, int64_t chunk, mbs_task_f task, mbs_round_f round, void *data);

/* Inside a task: suspend until the next round has run. */
void mbs_yield(void);
/* Whether the caller runs inside a task coroutine, and the pool thread it runs on (-1 outside). */
int mbs_in_task(void);
int mbs_tid(void);

typedef struct {
	int64_t rounds, yields, coroutines_peak, stacks;
	double round_s, task_s;          /* wall time in round callbacks, and in mbs_run overall */
	double phase_s;                  /* wall time of the compute phases (from a round's start to the last thread's arrival) */
	double busy_s;                   /* thread time spent running coroutines, summed over threads */
} mbs_stats_t;
void mbs_get_stats(const mbs_pool_t *p, mbs_stats_t *st);

#ifdef __cplusplus
}
#endif
#endif
