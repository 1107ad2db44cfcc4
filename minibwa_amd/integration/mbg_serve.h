/*
 * mbg_serve.h -- serving over the batched mapper: requests of read pairs or single reads arrive at
 * any time, are admitted continuously under a device-memory budget, and each request's SAM records
 * come back to its caller.
 *
 * Continuous batching: the mapper's worker pool runs one open run for the server's lifetime
 * (mbg_sched.h).  Whatever is queued when a front slot is free and the budget allows becomes a group
 * (mbg_map.h): its front pass (seeding, mb_anchor, the first mb_lchain_dp) runs on the device and one
 * task per read pair (or single read) joins the running pool, where it maps its reads (mb_map_sai),
 * pairs them (mb_pair) and writes the fragment's records (the output step's, map-main.c:207-240,
 * by mbg_fmt.c).  Requests admitted while others are mapping therefore map beside them: their
 * front passes overlap the earlier groups' host work, and all their DP and local-SW calls share the
 * device batches.  A request is done when its last fragment is written.  Pairs are mapped with
 * predefined insert-size statistics (MB_F_PE_PREDEF, as `minibwa map -I`), so a request's records
 * do not depend on the other requests in flight; single reads (long reads) form groups of their own,
 * since MB_F_PE applies to a whole group.  Reads carry no qualities, so QUAL is "*" as for FASTA
 * input.
 *
 * Admission: a group takes queued requests of the oldest request's kind, oldest first, while the
 * bases in flight stay within what the device-memory budget allows, and at most MBX_FRONT_SLOTS groups
 * are in flight.  The executor's device work buffers (mbx_mem) grow to the largest load seen and are
 * kept, so a load up to the largest in-flight bases so far needs no new memory, and a larger one about
 * bytes_per_base for each extra base: the budget left over the current work buffers, divided by that
 * rate, is how far the in-flight bases may grow.  The rate starts at a conservative guess and takes
 * the largest growth per base observed after the first update (whose growth also holds the buffers
 * that do not scale with the load).  The oldest request always runs when nothing is in flight, alone
 * if it is larger than that.
 *
 * The server holds the worker pool for its lifetime: the CLI pipeline and mb_map_batch share that pool
 * and wait for mbg_server_stop.
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
	int32_t max_reqs;          /* most requests in one group; 0: no limit */
	int32_t solo;              /* every request alone: a group of its own, nothing else in flight (the reference
	                              point for batch invariance) */
} mbg_serve_opt_t;

typedef struct {
	int64_t batches, reqs, reads, bases;   /* batches: groups finished */
	int64_t max_batch_bases;   /* the most bases in flight at once */
	int64_t gated;             /* groups closed by the memory budget while requests of their kind waited */
	int64_t dev_budget, work_bytes;
	double bytes_per_base;     /* the admission rate now (device bytes per base beyond the in-flight peak) */
	double busy_s;             /* group lifetimes, front pass to last fragment, summed */
} mbg_serve_stats_t;

/* opt_pe: options for read pairs (MB_F_PE and MB_F_PE_PREDEF with pe_avg/std/lo/hi set); opt_se: for
 * single reads.  Both are copied.  Starts the admission thread and the worker pool's open run. */
mbg_server_t *mbg_server_start(const mb_opt_t *opt_pe, const mb_opt_t *opt_se, const mb_idx_t *idx, const mbg_serve_opt_t *so);

/* Queues a request: paired != 0: n reads forming n/2 pairs (read 2i and 2i+1 are mates), else n
 * single reads; names and sequences (ASCII) are copied.  Returns the request's id (0, 1, ... in
 * submission order), or -1 for an empty request or after mbg_server_stop. */
int64_t mbg_server_submit(mbg_server_t *s, int paired, int32_t n, const char *const *name, const char *const *seq);

/* Blocks until request id is done; *sam receives its records (malloc'd, the caller frees) and *len
 * their length; t_start / t_done (may be null) the monotonic times its group started and it was
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
