/*
 * mbg_sched.h -- a worker pool that runs the mapper's per-read host code as coroutines and turns its
 * blocking per-call work into device batches.
 *
 * The reference maps one read at a time: mb_map_sai (map-algo.c:571-700) and mb_pair (pe.c:448-582)
 * call the DP and local-SW kernels one job at a time, each call depending on the previous one's
 * result (align.c:548-757: the left extension, the gap fills with their z-drop break and second pass,
 * the right extension; pe.c:319-378: the rescue candidates).  Here every read (or read pair) runs in a
 * coroutine on one of n host threads.  A hook that needs the device calls mbs_wait() with its request:
 * the coroutine is suspended and its thread goes on with other coroutines and new tasks.  A dispatcher
 * thread collects the posted requests and hands them to the batch callback as one device batch -- as
 * soon as enough are pending, or when no worker has anything left to run -- and each finished request's
 * coroutine is resumed by the thread that started it.  A read's own calls keep their order and their
 * data dependences exactly, while the device sees the pending calls of many reads at once, and the
 * host threads keep working while a batch is on the device.
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

/* Runs task i; may call mbs_wait() any number of times. tid is the pool thread (0..n_threads-1). */
typedef void (*mbs_task_f)(void *data, int64_t i, int tid);
/* Runs one batch of posted requests (on the dispatcher thread); returns when all are done. */
typedef void (*mbs_batch_f)(void *data, void **req, int64_t n);

mbs_pool_t *mbs_pool_init(int n_threads, int64_t stack_bytes);
void mbs_pool_destroy(mbs_pool_t *p);
int mbs_pool_size(const mbs_pool_t *p);

/* Runs task(data, i, tid) for every i in [0, n) and returns when all have finished.  Tasks are taken
 * in chunks of `chunk` consecutive indices.  batch may be null when no task waits.  flush: the number
 * of pending requests that starts a batch while workers still have work (0: a default). */
void mbs_run(mbs_pool_t *p, int64_t n, int64_t chunk, mbs_task_f task, mbs_batch_f batch, void *data, int64_t flush);

/* An open run: mbs_open starts a run of no tasks and returns; mbs_add(p, n) makes tasks [k, k + n)
 * runnable, where k (returned) is the number of tasks added before, and may be called from any thread
 * while the run is open (from inside a task too); mbs_close waits until every added task has finished
 * and ends the run.  One run (mbs_run or open) per pool at a time. */
void mbs_open(mbs_pool_t *p, int64_t chunk, mbs_task_f task, mbs_batch_f batch, void *data, int64_t flush);
int64_t mbs_add(mbs_pool_t *p, int64_t n);
void mbs_close(mbs_pool_t *p);

/* Inside a task: post req and suspend until the batch holding it has run. */
void mbs_wait(void *req);
/* Whether the caller runs inside a task coroutine, and the pool thread it runs on (-1 outside). */
int mbs_in_task(void);
int mbs_tid(void);

typedef struct {
	int64_t batches, waits, coroutines_peak, stacks;
	double batch_s;                  /* dispatcher time inside the batch callback */
	double task_s;                   /* wall time of mbs_run calls */
	double busy_s;                   /* thread time spent running coroutines, summed over threads */
	double idle_s;                   /* thread time spent waiting for finished requests, summed over threads */
} mbs_stats_t;
void mbs_get_stats(const mbs_pool_t *p, mbs_stats_t *st);

#ifdef __cplusplus
}
#endif
#endif
