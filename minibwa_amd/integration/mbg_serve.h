/*
 * mbg_serve.h -- serving over the batched mapper: requests of read pairs or single reads arrive at
 * any time, are admitted into device batches continuously (whatever is queued when the previous
 * batch finishes forms the next one) under a device-memory budget, and each request's SAM records
 * come back to its caller.
 *
 * A batch is one mb_map_batch call (map-algo.c:703-770) of the driver in mbg_map.c: one front pass
 * on the device, the mapping coroutines with their device DP rounds, and pairing; a second thread
 * writes a mapped batch's records (requests in parallel) while the next batch maps.  Pairs are mapped
 * with predefined insert-size statistics (MB_F_PE_PREDEF, as `minibwa map -I`), so that a request's
 * records do not depend on the other requests in its batch; single reads (long reads) form batches
 * of their own, since MB_F_PE applies to a whole batch.  The records are the output step's
 * (map-main.c:207-240), written by mbg_fmt.c; reads carry no qualities, so QUAL is "*" as for FASTA
 * input.
 *
 * Admission: a batch takes queued requests of the oldest request's kind, oldest first, while its
 * bases stay within what the device-memory budget allows.  The executor's device work buffers
 * (mbx_mem) grow to the largest batch seen and are kept, so a batch up to that size needs no new
 * memory, and a larger one needs about bytes_per_base for each extra base: the budget left over the
 * current work buffers, divided by that rate, is how far the largest batch may grow.  The rate starts
 * at a conservative guess and takes the largest growth per base observed after the first batch (whose
 * growth also holds the buffers that do not scale with the batch).  The oldest request always runs,
 * alone if it is larger than that.
 */
#ifndef MBG_SERVE_H
#define MBG_SERVE_H

#include <stdint.h>
#include "mbpriv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbg_server_s mbg_server_t;

typedef struct {
	int64_t dev_budget;        /* device bytes the work buffers may reach; 0: what they hold at start + half the free memory */
	double bytes_per_base0;    /* starting estimate of device bytes per read base (0: 4096) */
	int32_t max_reqs;          /* most requests in one batch; 0: no limit */
	int32_t solo;              /* every request in a batch of its own (the reference point for batch invariance) */
} mbg_serve_opt_t;

typedef struct {
	int64_t batches, reqs, reads, bases;
	int64_t max_batch_bases;   /* the largest batch admitted */
	int64_t gated;             /* batches closed by the memory budget while requests of their kind waited */
	int64_t dev_budget, work_bytes;
	double bytes_per_base;     /* the admission rate now (device bytes per base beyond the largest batch) */
	double busy_s;             /* time inside batches */
} mbg_serve_stats_t;

/* opt_pe: options for read pairs (MB_F_PE and MB_F_PE_PREDEF with pe_avg/std/lo/hi set); opt_se: for
 * single reads.  Both are copied.  Starts the batching thread. */
mbg_server_t *mbg_server_start(const mb_opt_t *opt_pe, const mb_opt_t *opt_se, const mb_idx_t *idx, const mbg_serve_opt_t *so);

/* Queues a request: paired != 0: n reads forming n/2 pairs (read 2i and 2i+1 are mates), else n
 * single reads; names and sequences (ASCII) are copied.  Returns the request's id (0, 1, ... in
 * submission order), or -1 for an empty request or after mbg_server_stop. */
int64_t mbg_server_submit(mbg_server_t *s, int paired, int32_t n, const char *const *name, const char *const *seq);

/* Blocks until request id is done; *sam receives its records (malloc'd, the caller frees) and *len
 * their length; t_start / t_done (may be null) the monotonic times its batch started and it was
 * done.  Returns 0, or -1 for an unknown id or one already collected. */
int mbg_server_wait(mbg_server_t *s, int64_t id, char **sam, int64_t *len, double *t_start, double *t_done);

/* Finishes the queued requests, then stops the thread and frees the server (uncollected results too). */
void mbg_server_stop(mbg_server_t *s);

void mbg_server_stats(mbg_server_t *s, mbg_serve_stats_t *st);

/* the monotonic clock the times above use */
double mbg_server_now(void);

#ifdef __cplusplus
}
#endif
#endif
