/*
 * map_exec_cpu.c -- test infrastructure: the batched mapper's device interface (mbg_exec.h) answered
 * by the reference's own functions on the CPU, so that the driver (minibwa_amd/integration/mbg_map.c:
 * its coroutine rounds, hooks, front-pass plumbing and pipeline step) can be checked against the
 * reference binary's SAM without a GPU.  Linked only into oracle/_ref/minibwa-batched-cpu by
 * oracle/Makefile; the product links mbg_exec_gpu.c instead.
 *
 *   front   __real_mb_seed_intv_batch (seed.c:56-103), then per read __real_mb_anchor
 *           (seed.c:268-339) and __real_mb_lchain_dp (lchain.c:145-231) with the read's options
 *   DP      __real_ksw_extd2_sse / __real_ksw_extz2_sse (ksw2.h:73-80)
 *   ll      __real_ksw_ll_qinit + __real_ksw_ll_u8_core / __real_ksw_ll_i16_core (ksw2.h:101-103)
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "kalloc.h"
#include "mbpriv.h"
#include "ksw2.h"
#include "mbg_exec.h"

void __real_mb_seed_intv_batch(void *km, const mb_bwt_t *bwt, int32_t n_seq, const int32_t *len, uint8_t *const *seq, int32_t min_len,
                               int32_t max_sub_occ, mb_sai_v *v);
void __real_mb_bwt_sa_batch(void *km, const mb_bwt_t *bwt, int64_t n, uint64_t *x);
double __real_mb_anchor(void *km, const mb_idx_t *idx, mb_sai_v *u, int32_t min_len, int32_t qlen, const uint8_t *qseq,
                        l2b_meth_t mt, int32_t max_occ, mb_anchor_v *v);
mb_anchor_t *__real_mb_lchain_dp(void *km, const l2b_t *l2b, int max_dist_x, int max_dist_y, int bw, int max_skip, int max_iter,
                                 int min_sc, float chn_pen_gap, int64_t n, mb_anchor_t *a, int *n_u_, uint64_t **_u);
void __real_ksw_extd2_sse(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                          int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez);
void __real_ksw_extz2_sse(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                          int8_t q, int8_t e, int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez);
void *__real_ksw_ll_qinit(void *km, int size, int qlen, const uint8_t *query, int m, const int8_t *mat);
ksw_llrst_t __real_ksw_ll_u8_core(void *q, int tlen, const uint8_t *target, int gapo, int gape, int xtra);
ksw_llrst_t __real_ksw_ll_i16_core(void *q, int tlen, const uint8_t *target, int gapo, int gape, int xtra);

static const mb_idx_t *g_idx;
static mbx_stats_t g_st;
static pthread_mutex_t g_st_mu = PTHREAD_MUTEX_INITIALIZER;

/* front results, one set per slot */
typedef struct {
	int64_t *intv_off, *n_intv, *anc_off, *n_anc, *n_chn;
	int32_t *n_u;
	mb_sai_t *intv;
	mb_anchor_t *anc, *chn;
	uint64_t *u;
} fres_t;
static fres_t g_fr[MBX_FRONT_SLOTS];

const char *mbx_name(void) { return "cpu-reference"; }

void mbx_index(const mb_idx_t *idx) { g_idx = idx; }

static void *xrealloc(void *p, size_t n)
{
	p = realloc(p, n ? n : 1);
	if (p == 0) { fprintf(stderr, "[map_exec_cpu] out of memory\n"); abort(); }
	return p;
}

void mbx_front(const mbx_front_in_t *in, mbx_front_t *out)
{
	int32_t n = in->n, r;
	uint8_t **seq = (uint8_t**)xrealloc(0, (n + 1) * sizeof(uint8_t*));
	mb_sai_v *sai = (mb_sai_v*)calloc(n + 1, sizeof(mb_sai_v));
	int64_t ti = 0, ta = 0;
	fres_t *F = &g_fr[in->slot >= 0 && in->slot < MBX_FRONT_SLOTS ? in->slot : 0];
	for (r = 0; r < n; ++r) seq[r] = (uint8_t*)in->nt4 + in->off[r];
	__real_mb_seed_intv_batch(0, g_idx->bwt, n, in->len, seq, in->min_len, in->max_sub_occ, sai);
	F->intv_off = (int64_t*)xrealloc(F->intv_off, (n + 1) * 8); F->n_intv = (int64_t*)xrealloc(F->n_intv, (n + 1) * 8);
	F->anc_off = (int64_t*)xrealloc(F->anc_off, (n + 1) * 8); F->n_anc = (int64_t*)xrealloc(F->n_anc, (n + 1) * 8);
	F->n_chn = (int64_t*)xrealloc(F->n_chn, (n + 1) * 8); F->n_u = (int32_t*)xrealloc(F->n_u, (n + 1) * 4);
	for (r = 0; r < n; ++r) {
		int64_t k, s = 0;
		for (k = 0; k < (int64_t)sai[r].n; ++k) s += sai[r].a[k].size < (uint64_t)in->max_occ ? (int64_t)sai[r].a[k].size : in->max_occ;
		F->intv_off[r] = ti, ti += sai[r].n;
		F->anc_off[r] = ta, ta += s;
	}
	F->intv_off[n] = ti, F->anc_off[n] = ta;
	F->intv = (mb_sai_t*)xrealloc(F->intv, ti * sizeof(mb_sai_t));
	F->anc = (mb_anchor_t*)xrealloc(F->anc, ta * sizeof(mb_anchor_t));
	F->chn = (mb_anchor_t*)xrealloc(F->chn, ta * sizeof(mb_anchor_t));
	F->u = (uint64_t*)xrealloc(F->u, ta * sizeof(uint64_t));
	for (r = 0; r < n; ++r) {
		mb_anchor_v v = {0, 0, 0};
		F->n_anc[r] = -1, F->n_u[r] = 0, F->n_chn[r] = 0;
		if (!(in->no_chain && in->no_chain[r]) && sai[r].n > 0) {
			const mbg_chain_opt_t *o = &in->copt[in->copt_idx[r]];
			mb_anchor_t *c;
			uint64_t *u = 0;
			int n_u = 0, i;
			int64_t n_c = 0;
			__real_mb_anchor(0, g_idx, &sai[r], in->min_len, in->len[r], seq[r], L2B_METH_NONE, in->max_occ, &v);
			F->n_anc[r] = v.n;
			memcpy(F->anc + F->anc_off[r], v.a, v.n * sizeof(mb_anchor_t));
			c = __real_mb_lchain_dp(0, g_idx->l2b, o->max_dist_x, o->max_dist_y, o->bw, o->max_skip, o->max_iter, o->min_sc, o->pen_gap,
			                        v.n, v.a, &n_u, &u);
			for (i = 0; i < n_u; ++i) n_c += (int32_t)u[i];
			if (c) {
				memcpy(F->u + F->anc_off[r], u, n_u * sizeof(uint64_t));
				memcpy(F->chn + F->anc_off[r], c, n_c * sizeof(mb_anchor_t));
				F->n_u[r] = n_u, F->n_chn[r] = n_c;
			}
			kfree(0, u); kfree(0, c);
		} else if (sai[r].n == 0) F->n_anc[r] = 0;
		F->n_intv[r] = sai[r].n;
		memcpy(F->intv + F->intv_off[r], sai[r].a, sai[r].n * sizeof(mb_sai_t));
		kfree(0, sai[r].a);
	}
	free(sai); free(seq);
	out->n = n;
	out->intv_off = F->intv_off, out->n_intv = F->n_intv, out->intv = F->intv;
	out->anc_off = F->anc_off, out->n_anc = F->n_anc, out->anc = F->anc;
	out->n_u = F->n_u, out->u = F->u, out->n_chn = F->n_chn, out->chn = F->chn;
	out->dev_ms = 0;
	pthread_mutex_lock(&g_st_mu);
	++g_st.fronts, g_st.front_reads += n;
	pthread_mutex_unlock(&g_st_mu);
}

void mbx_run_dp(int dual, int32_t n, mbx_req_t *const *req)
{
	int32_t i;
	ksw_extz_t ez;
	memset(&ez, 0, sizeof(ez));
	for (i = 0; i < n; ++i) {
		mbx_req_t *q = req[i];
		if (dual) __real_ksw_extd2_sse(0, q->qlen, q->q, q->tlen, q->t, q->m, q->mat, q->gq, q->ge, q->gq2, q->ge2, q->w, q->zdrop,
		                               q->end_bonus, q->flag, &ez);
		else __real_ksw_extz2_sse(0, q->qlen, q->q, q->tlen, q->t, q->m, q->mat, q->gq, q->ge, q->w, q->zdrop, q->end_bonus, q->flag, &ez);
		q->dp.max = ez.max, q->dp.zdropped = ez.zdropped, q->dp.max_q = ez.max_q, q->dp.max_t = ez.max_t;
		q->dp.mqe = ez.mqe, q->dp.mqe_t = ez.mqe_t, q->dp.mte = ez.mte, q->dp.mte_q = ez.mte_q;
		q->dp.score = ez.score, q->dp.reach_end = ez.reach_end, q->dp.n_cigar = ez.n_cigar, q->dp.cells = 0;
		if (ez.n_cigar > q->qlen + q->tlen + 1) { fprintf(stderr, "[map_exec_cpu] CIGAR longer than its room\n"); abort(); }
		memcpy(q->cigar, ez.cigar, (size_t)ez.n_cigar * 4);
	}
	kfree(0, ez.cigar);
	pthread_mutex_lock(&g_st_mu);
	++g_st.dp_calls, g_st.dp_jobs += n;
	pthread_mutex_unlock(&g_st_mu);
}

void mbx_run_ll(int32_t n, mbx_req_t *const *req)
{
	int32_t i;
	for (i = 0; i < n; ++i) {
		mbx_req_t *q = req[i];
		void *qp = __real_ksw_ll_qinit(0, q->size, q->qlen, q->q, q->m, q->mat);
		ksw_llrst_t r = q->size == 1 ? __real_ksw_ll_u8_core(qp, q->tlen, q->t, q->gapo, q->gape, q->xtra)
		                             : __real_ksw_ll_i16_core(qp, q->tlen, q->t, q->gapo, q->gape, q->xtra);
		kfree(0, qp);
		q->ll.score = r.score, q->ll.te = r.te, q->ll.qe = r.qe, q->ll.score2 = r.score2, q->ll.te2 = r.te2;
	}
	pthread_mutex_lock(&g_st_mu);
	++g_st.ll_calls, g_st.ll_jobs += n;
	pthread_mutex_unlock(&g_st_mu);
}

void mbx_sa(int64_t n, uint64_t *x) { __real_mb_bwt_sa_batch(0, g_idx->bwt, n, x); }

void mbx_seed(void *km, int32_t n, const int32_t *len, uint8_t *const *seq, int32_t min_len, int32_t max_sub_occ, mb_sai_v *v)
{
	__real_mb_seed_intv_batch(km, g_idx->bwt, n, len, seq, min_len, max_sub_occ, v);
}

void mbx_get_stats(mbx_stats_t *st) { *st = g_st; memset(&st->dev, 0, sizeof(st->dev)); }

void mbx_mem(int64_t *work, int64_t *dev_free, int64_t *dev_total)
{
	*work = *dev_free = *dev_total = 0;
}
