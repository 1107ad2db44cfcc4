/*
 * mbg_exec.h -- what the batched mapper (mbg_map.c) asks of the device, as one small interface.
 *
 * The product links mbg_exec_gpu.c, which runs every request through libmbg's C ABI (include/mbg.h).
 * The CPU check of the batching machinery links oracle/map_exec_cpu.c instead, which answers the
 * same requests with the reference's own functions, so that the driver's rounds, hooks and output
 * can be checked against the reference binary on a machine without a GPU.
 */
#ifndef MBG_EXEC_H
#define MBG_EXEC_H

#include <stdint.h>
#include "mbpriv.h"
#include "mbg.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { MBX_D2 = 1, MBX_Z2 = 2, MBX_LL = 3 };

/* One deferred call: ksw_extd2_sse / ksw_extz2_sse (ksw2.h:73-80) or ksw_ll_qinit + ksw_ll_u8_core /
 * ksw_ll_i16_core (ksw2.h:101-103).  Sequences and the matrix stay where the caller keeps them (the
 * suspended coroutine's memory) until the round has run. */
typedef struct {
	int32_t kind;
	int32_t qlen, tlen;
	const uint8_t *q, *t;
	int8_t m, gq, ge, gq2, ge2;
	int8_t mat[25];
	int32_t w, zdrop, end_bonus, flag;     /* DP */
	int32_t size, gapo, gape, xtra;        /* local SW: size 1 = u8 core, 2 = i16 core */
	uint32_t *cigar;                       /* DP: room for qlen + tlen operations, the poster's own memory */
	/* results, filled by the batch */
	mbg_dp_result_t dp;                    /* its CIGAR (dp.n_cigar operations) is written to cigar */
	mbg_ll_result_t ll;
} mbx_req_t;

/* One mini-batch's front half of mb_map_sai (map-algo.c:588-631) for reads without a methylation
 * conversion: two-pass seeding (mb_seed_intv_batch), mb_anchor and the first mb_lchain_dp call, with
 * each read's own chaining options (copt[copt_idx[r]], from mb_opt_adap for its length). */
typedef struct {
	int32_t slot;                  /* 0..MBX_FRONT_SLOTS-1: results of different slots stay valid together */
	int32_t n;
	const uint8_t *nt4;            /* read r: nt4 + off[r], len[r] bases (already converted for methylation) */
	const int64_t *off;
	const int32_t *len;
	const uint8_t *no_chain;       /* read r != 0: seeding only (a methylation-converted read); may be null */
	int32_t min_len, max_sub_occ, max_occ;
	int32_t n_copt;
	const mbg_chain_opt_t *copt;
	const int32_t *copt_idx;
} mbx_front_in_t;

/* Per read r: its intervals (deduplicated, any order), anchors in mb_anchor's order, and chains with
 * their anchors as mb_lchain_dp returns them.  Valid until the next mbx_front call on the same slot;
 * calls on different slots may run concurrently (the CLI's three pipeline steps hold up to three
 * mini-batches). */
#define MBX_FRONT_SLOTS 6
typedef struct {
	int32_t n;
	const int64_t *intv_off, *n_intv;
	const mb_sai_t *intv;
	const int64_t *anc_off, *n_anc;        /* n_anc[r] < 0: not computed (seeding only) */
	const mb_anchor_t *anc;
	const int32_t *n_u;                    /* chains at u + anc_off[r] */
	const uint64_t *u;
	const int64_t *n_chn;                  /* chained anchors at chn + anc_off[r] */
	const mb_anchor_t *chn;
	double dev_ms;                         /* device time of the front (0 if not measured) */
} mbx_front_t;

const char *mbx_name(void);
/* The index main_map loaded goes where the executor needs it (HBM for the GPU). */
void mbx_index(const mb_idx_t *idx);
void mbx_front(const mbx_front_in_t *in, mbx_front_t *out);
/* Runs n requests of one kind (MBX_D2 and MBX_Z2: dual = kind == MBX_D2; MBX_LL). */
void mbx_run_dp(int dual, int32_t n, mbx_req_t *const *r);
void mbx_run_ll(int32_t n, mbx_req_t *const *r);
/* mb_bwt_sa_batch and mb_seed_intv_batch semantics, for calls outside the batched driver. */
void mbx_sa(int64_t n, uint64_t *x);
void mbx_seed(void *km, int32_t n, const int32_t *len, uint8_t *const *seq, int32_t min_len, int32_t max_sub_occ, mb_sai_v *v);
/* Device memory: the executor's work buffers (high-water, all contexts), and the device's free and
 * total memory; all zero for an executor without a device. */
void mbx_mem(int64_t *work, int64_t *dev_free, int64_t *dev_total);
/* Accumulated device-side counters for the run summary (may be all zero). */
typedef struct {
	int64_t dp_calls, dp_jobs, ll_calls, ll_jobs, fronts, front_reads;
	double dp_s, ll_s, front_s;            /* host wall time inside the executor's calls */
	double dp_pack_s;                      /* of dp_s: packing the rounds' sequences */
	mbg_counters_t dev;                    /* the device's own counters (zero for a CPU executor) */
} mbx_stats_t;
void mbx_get_stats(mbx_stats_t *st);

#ifdef __cplusplus
}
#endif
#endif
