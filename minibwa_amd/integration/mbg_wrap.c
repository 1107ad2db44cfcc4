/*
 * mbg_wrap.c -- binds the reference's hot-path hooks to libmbg without touching its sources.
 *
 * Linked into the unmodified minibwa objects with
 *   -Wl,--wrap=mb_seed_intv_batch -Wl,--wrap=mb_bwt_sa_batch -Wl,--wrap=ksw_extd2_sse -Wl,--wrap=ksw_extz2_sse
 *   -Wl,--wrap=ksw_ll_qinit -Wl,--wrap=ksw_ll_u8_core -Wl,--wrap=ksw_ll_i16_core -Wl,--wrap=ksw_ll_i16
 *   -Wl,--wrap=mb_anchor -Wl,--wrap=mb_lchain_dp
 * every call site of those functions (map-main.c:75, map-algo.c:733, seed.c:180, align.c:374-376,
 * pe.c:267-272,287, map-algo.c:610,614,634) reaches the __wrap_ definitions below, which run the work on the GPU through the
 * C ABI of include/mbg.h and hand results back in the reference's own containers:
 *
 *   mb_seed_intv_batch (mbpriv.h:57, seed.c:56-103)   -> mbg_seed_batch
 *   mb_bwt_sa_batch    (bwt.h:73, bwt.c:511-541)       -> mbg_sa_batch
 *   ksw_extd2_sse      (ksw2.h:79-80)                  -> mbg_extd2_batch (one job)
 *   ksw_extz2_sse      (ksw2.h:73-74)                  -> mbg_extz2_batch (one job)
 *   ksw_ll_qinit + ksw_ll_u8_core / ksw_ll_i16_core (ksw2.h:101-103)
 *                                                      -> mbg_ll_batch (one job)
 *   mb_anchor          (mbpriv.h:58, seed.c:268-339)   -> mbg_anchor_batch (one read; reads with a
 *                                                         methylation conversion keep the reference's
 *                                                         own, which then locates through mbg_sa_batch)
 *   mb_lchain_dp       (mbpriv.h:62-63, lchain.c:145-240) -> mbg_lchain_batch (one read)
 *
 * MBG_WRAP_STATS=1 prints the number of calls each hook took, at exit, to stderr.
 * The index is uploaded once, on the first hooked call, from the mb_bwt_t the caller passes (the
 * k-mer cache is recomputed on the device with the caller's pre_len).  MBG_DEVICE selects the HIP
 * device (default 0).  Any library error aborts, as the reference does on internal failures.
 *
 * Containers follow the reference's allocation rules: seed intervals and anchors are grown with
 * kalloc's krealloc on the caller's km (Kgrow, kalloc.h:34-41; map-algo.c and lchain.c release them
 * with kalloc's kfree), CIGAR storage and the ksw_ll query record with the allocator ksw2.h selects
 * for this build (ksw2.h:113-121: plain malloc/realloc/free unless HAVE_KALLOC, as align.c and pe.c
 * see it), and ksw_extz_t is filled field for field (ksw2.h:36-45).  ksw2.h's macros shadow
 * kalloc's names in this file, so kalloc's own functions are called as (krealloc)(...).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "kalloc.h"
#include "bwt.h"
#include "mbpriv.h"
#include "ksw2.h"
#include "mbg.h"

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static mbg_ctx_t *g_ctx;
static const mb_bwt_t *g_bwt;
static const l2b_t *g_l2b;
static long g_calls[7];                    /* seed, sa, extd2, extz2, ll, anchor, chain */

static void report(void)
{
	fprintf(stderr, "[mbg_wrap] calls: seed %ld sa %ld anchor %ld chain %ld extd2 %ld extz2 %ld ll %ld\n", g_calls[0],
	        g_calls[1], g_calls[5], g_calls[6], g_calls[2], g_calls[3], g_calls[4]);
}

static void count(int k)
{
	pthread_mutex_lock(&g_lock);
	++g_calls[k];
	pthread_mutex_unlock(&g_lock);
}

static void die(const char *what, int rc)
{
	fprintf(stderr, "[mbg_wrap] %s failed: %s (%s)\n", what, mbg_strerror(rc), mbg_last_hip_error());
	abort();
}

static mbg_ctx_t *ctx_for(const mb_bwt_t *bwt)
{
	pthread_mutex_lock(&g_lock);
	if (g_ctx == 0) {
		const char *d = getenv("MBG_DEVICE"), *st = getenv("MBG_WRAP_STATS");
		int rc = mbg_init(d ? atoi(d) : 0, &g_ctx);
		if (rc != MBG_OK) die("mbg_init", rc);
		if (st && atoi(st)) atexit(report);
	}
	if (bwt && bwt != g_bwt) {
		mbg_bwt_desc_t dsc;
		int rc;
		dsc.primary = bwt->primary;
		memcpy(dsc.L2, bwt->L2, sizeof(dsc.L2));
		dsc.seq_len = bwt->seq_len;
		dsc.data = bwt->data;
		dsc.sa_bit = bwt->sa_bit;
		dsc.n_sa = bwt->n_sa;
		dsc.sa = bwt->sa;
		rc = mbg_bwt_upload(g_ctx, &dsc, bwt->pre ? (int32_t)bwt->pre_len : 0);
		if (rc != MBG_OK) die("mbg_bwt_upload", rc);
		g_bwt = bwt;
	}
	pthread_mutex_unlock(&g_lock);
	return g_ctx;
}

void __wrap_mb_seed_intv_batch(void *km, const mb_bwt_t *bwt, int32_t n_seq, const int32_t *len, uint8_t *const *seq,
                               int32_t min_len, int32_t max_sub_occ, mb_sai_v *v)
{
	mbg_ctx_t *c = ctx_for(bwt);
	mbg_sai_t *out = 0;
	int64_t *off;
	int32_t i;
	int rc;
	if (n_seq <= 0) return;
	count(0);
	off = (int64_t*)malloc(sizeof(int64_t) * (n_seq + 1));
	rc = mbg_seed_batch(c, n_seq, len, (const uint8_t *const*)seq, min_len, max_sub_occ, &out, off);
	if (rc != MBG_OK) die("mbg_seed_batch", rc);
	for (i = 0; i < n_seq; ++i) {
		int64_t k, n = off[i + 1] - off[i];
		v[i].n = 0;
		if ((size_t)n > v[i].m) {
			v[i].m = (size_t)n + ((size_t)n >> 1) + 16;
			v[i].a = (mb_sai_t*)(krealloc)(km, v[i].a, v[i].m * sizeof(mb_sai_t));
		}
		for (k = 0; k < n; ++k) {
			const mbg_sai_t *s = &out[off[i] + k];
			mb_sai_t *d = &v[i].a[v[i].n++];
			d->x[0] = s->x[0], d->x[1] = s->x[1], d->size = s->size, d->info = s->info;
		}
	}
	free(out);
	free(off);
}

void __wrap_mb_bwt_sa_batch(void *km, const mb_bwt_t *bwt, int64_t n, uint64_t *x)
{
	int rc;
	(void)km;
	if (n <= 0) return;
	count(1);
	rc = mbg_sa_batch(ctx_for(bwt), n, x);
	if (rc != MBG_OK) die("mbg_sa_batch", rc);
}

static void dp_one(int dual, void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                   const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus,
                   int flag, ksw_extz_t *ez)
{
	mbg_ctx_t *c = ctx_for(0);
	mbg_dp_job_t j;
	mbg_dp_result_t r;
	uint32_t *cig;
	int64_t cap;
	int rc;
	count(dual ? 2 : 3);
	memset(&j, 0, sizeof(j));
	j.qlen = qlen > 0 ? qlen : 0;      /* the reference returns the reset state for qlen/tlen <= 0 */
	j.tlen = tlen > 0 ? tlen : 0;
	j.q = q, j.e = e, j.q2 = q2, j.e2 = e2;
	j.w = w, j.zdrop = zdrop, j.end_bonus = end_bonus, j.flag = flag;
	cap = (int64_t)j.qlen + j.tlen + 1;
	cig = (uint32_t*)malloc(cap * sizeof(uint32_t));
	rc = dual ? mbg_extd2_batch(c, 1, &j, query, j.qlen, target, j.tlen, m, mat, &r, cig, cap)
	          : mbg_extz2_batch(c, 1, &j, query, j.qlen, target, j.tlen, m, mat, &r, cig, cap);
	if (rc != MBG_OK) die(dual ? "mbg_extd2_batch" : "mbg_extz2_batch", rc);
	ez->max = (uint32_t)r.max & 0x7fffffff, ez->zdropped = r.zdropped ? 1 : 0;
	ez->max_q = r.max_q, ez->max_t = r.max_t;
	ez->mqe = r.mqe, ez->mqe_t = r.mqe_t;
	ez->mte = r.mte, ez->mte_q = r.mte_q;
	ez->score = r.score, ez->reach_end = r.reach_end;
	ez->n_cigar = r.n_cigar;
	if (r.n_cigar > ez->m_cigar) {
		ez->m_cigar = r.n_cigar;
		ez->cigar = (uint32_t*)krealloc(km, ez->cigar, (size_t)ez->m_cigar * 4);
	}
	if (r.n_cigar > 0) memcpy(ez->cigar, cig, (size_t)r.n_cigar * 4);
	free(cig);
}

void __wrap_ksw_extd2_sse(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                          const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus,
                          int flag, ksw_extz_t *ez)
{
	dp_one(1, km, qlen, query, tlen, target, m, mat, q, e, q2, e2, w, zdrop, end_bonus, flag, ez);
}

void __wrap_ksw_extz2_sse(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                          const int8_t *mat, int8_t q, int8_t e, int w, int zdrop, int end_bonus, int flag,
                          ksw_extz_t *ez)
{
	dp_one(0, km, qlen, query, tlen, target, m, mat, q, e, q, e, w, zdrop, end_bonus, flag, ez);
}

/* ksw_ll_qinit's query profile is kept as the query and matrix themselves; the core call runs the
 * job.  The record comes from the caller's km, so the caller's kfree (pe.c:281) releases it. */
typedef struct { int32_t qlen, m; int8_t mat[25]; uint8_t q[]; } ll_query_t;

void *__wrap_ksw_ll_qinit(void *km, int size, int qlen, const uint8_t *query, int m, const int8_t *mat)
{
	ll_query_t *q = (ll_query_t*)kmalloc(km, sizeof(ll_query_t) + (qlen > 0 ? qlen : 0) + 1);
	(void)size;                        /* the core the caller picks names the size */
	if (m < 1 || m > 5) die("ksw_ll_qinit (alphabet size)", MBG_EUNSUPPORTED);
	q->qlen = qlen > 0 ? qlen : 0, q->m = m;
	memcpy(q->mat, mat, (size_t)m * m);
	if (q->qlen) memcpy(q->q, query, q->qlen);
	return q;
}

static ksw_llrst_t ll_one(void *q_, int tlen, const uint8_t *target, int gapo, int gape, int xtra, int size)
{
	const ll_query_t *q = (const ll_query_t*)q_;
	mbg_ll_job_t j;
	mbg_ll_result_t r;
	ksw_llrst_t o;
	uint8_t dummy = 0;
	int rc;
	count(4);
	memset(&j, 0, sizeof(j));
	j.qlen = q->qlen, j.tlen = tlen > 0 ? tlen : 0, j.size = size;
	rc = mbg_ll_batch(ctx_for(0), 1, &j, q->qlen ? q->q : &dummy, q->qlen, j.tlen ? target : &dummy, j.tlen, (int8_t)q->m,
	                  q->mat, gapo, gape, xtra, &r);
	if (rc != MBG_OK) die("mbg_ll_batch", rc);
	o.score = r.score, o.te = r.te, o.qe = r.qe, o.score2 = r.score2, o.te2 = r.te2;
	return o;
}

ksw_llrst_t __wrap_ksw_ll_u8_core(void *q, int tlen, const uint8_t *target, int gapo, int gape, int xtra)
{
	return ll_one(q, tlen, target, gapo, gape, xtra, 1);
}

ksw_llrst_t __wrap_ksw_ll_i16_core(void *q, int tlen, const uint8_t *target, int gapo, int gape, int xtra)
{
	return ll_one(q, tlen, target, gapo, gape, xtra, 2);
}

/* ksw_ll_i16 (ksw2_ll_sse.c:335-341, called by mb_align1_inv, align.c:798) reaches the i16 core inside
 * its own file, past the hook above, with the query record __wrap_ksw_ll_qinit made: bound here too */
int __wrap_ksw_ll_i16(void *q, int tlen, const uint8_t *target, int gapo, int gape, int *qe, int *te)
{
	ksw_llrst_t r = ll_one(q, tlen, target, gapo, gape, 0, 2);
	*qe = r.qe, *te = r.te;
	return r.score;
}

/* the contig table goes to the device once per l2b */
static void l2b_for(mbg_ctx_t *c, const l2b_t *l)
{
	pthread_mutex_lock(&g_lock);
	if (l != g_l2b) {
		uint64_t *len = (uint64_t*)malloc(sizeof(uint64_t) * (l->n_ctg ? l->n_ctg : 1));
		mbg_l2b_desc_t d;
		uint64_t i;
		int rc;
		for (i = 0; i < l->n_ctg; ++i) len[i] = l->ctg[i].len;
		d.tot_len = l->tot_len, d.n_ctg = l->n_ctg, d.ctg_len = len;
		d.n_pac = l->n_pac, d.pac = l->pac, d.n_ambi = l->n_ambi, d.ambi = (const uint64_t*)l->ambi;
		rc = mbg_l2b_upload(c, &d);
		free(len);
		if (rc != MBG_OK) die("mbg_l2b_upload", rc);
		g_l2b = l;
	}
	pthread_mutex_unlock(&g_lock);
}

/* mb_anchor: the read's intervals and anchors cross the ABI as they are (mb_sai_t = mbg_sai_t,
 * mb_anchor_t = mbg_anchor_t, whose flag word holds flag:31 and flt:1); the anchor array is grown on
 * the caller's km to the reference's preallocation (seed.c:292-293).  The contig table goes to the
 * device with the first call on an index. */
double __real_mb_anchor(void *km, const mb_idx_t *idx, mb_sai_v *u, int32_t min_len, int32_t qlen, const uint8_t *qseq,
                        l2b_meth_t mt, int32_t max_occ, mb_anchor_v *v);

double __wrap_mb_anchor(void *km, const mb_idx_t *idx, mb_sai_v *u, int32_t min_len, int32_t qlen, const uint8_t *qseq,
                        l2b_meth_t mt, int32_t max_occ, mb_anchor_v *v)
{
	mbg_ctx_t *c;
	int64_t i, slots = 0, uoff[2], aoff[2], n_anc = 0, n_intv = 0;
	int rc;
	typedef char sai_layout[sizeof(mb_sai_t) == sizeof(mbg_sai_t) ? 1 : -1];
	typedef char anchor_layout[sizeof(mb_anchor_t) == sizeof(mbg_anchor_t) ? 1 : -1];
	(void)sizeof(sai_layout), (void)sizeof(anchor_layout);
	if (mt != L2B_METH_NONE) return __real_mb_anchor(km, idx, u, min_len, qlen, qseq, mt, max_occ, v);
	v->n = 0;
	if (u->n == 0) return 1.0;
	count(5);
	c = ctx_for(idx->bwt);
	l2b_for(c, idx->l2b);
	for (i = 0; i < (int64_t)u->n; ++i)
		slots += u->a[i].size < (uint64_t)max_occ ? (int64_t)u->a[i].size : max_occ;
	if (slots > v->m) {
		v->m = slots;
		v->a = (mb_anchor_t*)(krealloc)(km, v->a, (size_t)v->m * sizeof(mb_anchor_t));   /* kalloc's, not ksw2.h's macro */
	}
	uoff[0] = 0, uoff[1] = (int64_t)u->n, aoff[0] = 0, aoff[1] = slots;
	rc = mbg_anchor_batch(c, 1, (mbg_sai_t*)u->a, uoff, &qlen, max_occ, (mbg_anchor_t*)v->a, aoff, &n_anc, &n_intv);
	if (rc != MBG_OK) die("mbg_anchor_batch", rc);
	u->n = (size_t)n_intv;
	v->n = n_anc;
	return 1.0;
}

/* mb_lchain_dp: the input anchors are released and the chained anchors and chain list come from the
 * caller's km, with kalloc's own functions, as lchain.c does; no chain leaves both outputs empty. */
mb_anchor_t *__wrap_mb_lchain_dp(void *km, const l2b_t *l2b, int max_dist_x, int max_dist_y, int bw, int max_skip, int max_iter,
                                 int min_sc, float chn_pen_gap, int64_t n, mb_anchor_t *a, int *n_u_, uint64_t **_u)
{
	mbg_ctx_t *c;
	mbg_chain_opt_t o;
	int64_t off[2], n_out = 0;
	int32_t n_u = 0;
	uint64_t *u;
	mb_anchor_t *b;
	int rc;
	if (_u) *_u = 0, *n_u_ = 0;
	if (n == 0 || a == 0) {
		(kfree)(km, a);
		return 0;
	}
	if (n > INT32_MAX) die("mb_lchain_dp (more than 2^31 anchors)", MBG_EUNSUPPORTED);
	count(6);
	c = ctx_for(0);
	l2b_for(c, l2b);
	memset(&o, 0, sizeof(o));
	o.max_dist_x = max_dist_x, o.max_dist_y = max_dist_y, o.bw = bw, o.max_skip = max_skip, o.max_iter = max_iter;
	o.min_sc = min_sc, o.pen_gap = chn_pen_gap;
	u = (uint64_t*)(kmalloc)(km, (size_t)n * sizeof(uint64_t));
	b = (mb_anchor_t*)(kmalloc)(km, (size_t)n * sizeof(mb_anchor_t));
	off[0] = 0, off[1] = n;
	rc = mbg_lchain_batch(c, 1, (const mbg_anchor_t*)a, off, &n, &o, u, &n_u, (mbg_anchor_t*)b, &n_out);
	if (rc != MBG_OK) die("mbg_lchain_batch", rc);
	(kfree)(km, a);
	if (n_u == 0) {
		(kfree)(km, u); (kfree)(km, b);
		return 0;
	}
	if (_u) *_u = u, *n_u_ = n_u;
	else (kfree)(km, u);
	return b;
}
