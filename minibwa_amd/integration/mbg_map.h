/*
 * mbg_map.h -- groups of reads for the serving loop (mbg_serve.c), over the batched mapper's machinery
 * in mbg_map.c.
 *
 * A group is a set of reads (read pairs as reads 2i and 2i+1 when the options pair them) whose front
 * pass -- seeding, mb_anchor and the first mb_lchain_dp on the device -- runs when the group is made,
 * on one of the executor's front slots (results stay valid until the slot's next front).  Its reads are
 * then mapped by tasks of an open run of the worker pool (mbg_sched.h: mbs_open / mbs_add), one read or
 * one pair per task: mb_map_sai for each read and, for a pair, mb_pair with mb_map_batch's predefined
 * insert-size statistics (map-algo.c:755-766), so that a pair's hits depend on nothing outside it.
 * Tasks of several groups run in one open run, and their DP and local-SW calls share device batches.
 */
#ifndef MBG_MAP_H
#define MBG_MAP_H

#include <stdint.h>
#include "mbpriv.h"
#include "mbg_sched.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbg_group_s mbg_group_t;

/* n_seq reads (qlen, seq, name are kept by pointer until mbg_group_free); runs the front pass on
 * front slot `slot` (0..MBX_FRONT_SLOTS-1, not shared with another live group).  Null for no reads or
 * a methylation option without a methylation index. */
mbg_group_t *mbg_group_new(const mb_opt_t *opt, const mb_idx_t *idx, int32_t n_seq, const int32_t *qlen, const char *const *seq,
                           const char *const *name, int slot);
/* Inside a task of a run of mbg_map_pool()'s pool with mbg_map_round as its batch callback: maps read
 * r0 (n = 1), or reads r0 and r0 + 1 and pairs them (n = 2). */
void mbg_group_map(mbg_group_t *g, int32_t r0, int32_t n);
int32_t *mbg_group_n_hit(mbg_group_t *g);
mb_hit_t **mbg_group_hit(mbg_group_t *g);
void mbg_group_free(mbg_group_t *g);

/* the record formatter's set-up for a run with these options (the native formatter unless MBG_FMT=ref) */
void mbg_map_fmt_init(const mb_opt_t *opt);
mbs_pool_t *mbg_map_pool(int n_threads);
void mbg_map_round(void *data, void **reqs, int64_t n);
int64_t mbg_map_flush(void);

#ifdef __cplusplus
}
#endif
#endif
