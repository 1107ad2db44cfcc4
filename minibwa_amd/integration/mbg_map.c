/*
 * mbg_map.c -- the batched drop-in mapper: the reference CLI (`minibwa map`, `minibwa mem`) and its
 * public batch API (mb_map_batch, minibwa.h:170) with one device pass per mini-batch for seeding,
 * anchors and chains, and every DP and mate-rescue local-SW call of the mini-batch executed in device
 * rounds.  The reference's own host code -- option parsing, reading, mb_map_sai's hit bookkeeping,
 * mb_align1, pairing, rescue and SAM/PAF formatting -- runs unchanged; this file replaces only the
 * mapping step of its pipeline and binds the hook points below.
 *
 * Linked into the unmodified reference objects with -Wl,--wrap for:
 *   kt_pipeline          (kthread.h:10; only caller mb_map_file, map-main.c:292)  the mapping step of
 *                         worker_pipeline (map-main.c:176-192) is replaced by map_step() below; reading
 *                         (step 0) and output (step 2) stay the reference's
 *   mb_map_batch         (minibwa.h:170, map-algo.c:703-770)  the same machinery for API callers
 *   mb_idx_load          (minibwa.h:123, map-algo.c:19-42)  the loaded index is made resident
 *   mb_anchor, mb_lchain_dp (mbpriv.h:58,62)  answered from the mini-batch's device front pass
 *   ksw_extd2_sse, ksw_extz2_sse (ksw2.h:73-80)  deferred into the current round
 *   ksw_ll_qinit, ksw_ll_u8_core, ksw_ll_i16_core, ksw_ll_i16 (ksw2.h:101-104)  deferred likewise
 *   mb_seed_intv_batch, mb_seed_intv, mb_bwt_sa_batch  (mbpriv.h:56-57, bwt.h:73)  calls that reach
 *                         the hot path outside the batched driver run on the device one call at a time
 *
 * How a mini-batch is mapped (map_reads):
 *   1. front   every read's nt4 bases go to the device in one call: two-pass SMEM seeding, mb_anchor and
 *              the first mb_lchain_dp with the read's adaptive options (mbg_exec.h: mbx_front);
 *   2. SE      each read runs the reference's mb_map_sai in a coroutine (mbg_sched.h).  Its mb_anchor
 *              and mb_lchain_dp calls return the front's results; each DP call posts a request and
 *              yields; when all reads have yielded or finished, the round runs every posted DP job as
 *              one device batch, and the reads resume -- so the k-th DP call of every read runs in
 *              round k, and a read's calls keep their order and their dependences (the gap-fill loop's
 *              z-drop break and second pass, align.c:689-721, see the previous call's result);
 *   3. PE      the insert-size model (mb_pestat, or the predefined one) exactly as worker_pipeline
 *              chooses it (map-main.c:180-190), then mb_pair per pair in coroutines: each rescue
 *              candidate's ksw_ll and extension DP are deferred into rounds the same way.
 * Results are the reference's by construction: the same functions run on the same data, and every
 * deferred call returns what the reference's call returns (the device kernels are bit-exact).
 *
 * The layouts of pipeline_t and step_t (map-main.c:16-34, minibwa 0.7-r421) are restated below because
 * kt_pipeline hands them over as void*; map_step checks that the pipeline's fields agree with its
 * options before using them.
 *
 * Environment: MBG_MAP_THREADS (pool size for mb_map_batch; the CLI uses -t), MBG_MAP_STATS=1 (per
 * mini-batch timing on stderr).
 */
#define _GNU_SOURCE
#include <assert.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "kalloc.h"
#include "bwt.h"
#include "bseq.h"
#include "mbpriv.h"
#include "kthread.h"
#include "ksw2.h"
#include "mbg.h"
#include "mbg_exec.h"
#include "mbg_sched.h"

/* ---- the reference's pipeline records (map-main.c:16-34) --------------------------------------- */

typedef struct {
	int32_t n_fp, n_threads;
	int64_t n_base, n_seq;
	int64_t mb_size;
	const mb_opt_t *opt;
	mb_bseq_file_t **fp;
	const mb_idx_t *idx;
	FILE *fp_out;
} ref_pipeline_t;

typedef struct {
	const ref_pipeline_t *p;
	int32_t n_seq, n_frag, n_sb, n_pe;
	int32_t *n_hit, *seg_off, *seg_cnt;
	int32_t *sb_off, *sb_cnt;
	mb_pestat_t pes[4];
	mb_bseq1_t *seq;
	mb_hit_t **hit;
	mb_tbuf_t **tbuf;
} ref_step_t;

/* ---- global state ------------------------------------------------------------------------------ */

#define MAX_THREADS 256

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;      /* pool creation, calls outside rounds */
static mbs_pool_t *g_pool;
static mb_tbuf_t *g_tbuf[MAX_THREADS];
static const mb_idx_t *g_idx;
static int g_stats;

/* requests posted by the coroutines of each pool thread during the current round */
typedef struct { mbx_req_t **a; int64_t n, m; } req_list_t;
static req_list_t g_posted[MAX_THREADS];

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void die(const char *msg)
{
	fprintf(stderr, "[mbg_map] %s\n", msg);
	abort();
}

static void report(void)
{
	mbs_stats_t ss;
	mbx_stats_t xs;
	memset(&ss, 0, sizeof(ss));
	if (g_pool) mbs_get_stats(g_pool, &ss);
	mbx_get_stats(&xs);
	fprintf(stderr, "[mbg_map] summary: executor %s fronts %ld reads %ld front_s %.3f | rounds %ld yields %ld peak_coroutines %ld "
	        "round_s %.3f run_s %.3f | dp_calls %ld dp_jobs %ld dp_s %.3f | ll_calls %ld ll_jobs %ld ll_s %.3f\n",
	        mbx_name(), (long)xs.fronts, (long)xs.front_reads, xs.front_s, (long)ss.rounds, (long)ss.yields, (long)ss.coroutines_peak,
	        ss.round_s, ss.task_s, (long)xs.dp_calls, (long)xs.dp_jobs, xs.dp_s, (long)xs.ll_calls, (long)xs.ll_jobs, xs.ll_s);
}

static mbs_pool_t *pool_for(int n_threads)
{
	int i;
	pthread_mutex_lock(&g_mu);
	if (n_threads < 1) n_threads = 1;
	if (n_threads > MAX_THREADS) n_threads = MAX_THREADS;
	if (g_pool && mbs_pool_size(g_pool) != n_threads) {
		mbs_pool_destroy(g_pool);
		g_pool = 0;
	}
	if (g_pool == 0) {
		const char *s = getenv("MBG_MAP_STATS");
		if (s && atoi(s) && !g_stats) atexit(report);
		g_stats = s && atoi(s);
		g_pool = mbs_pool_init(n_threads, 256 * 1024);
		if (g_pool == 0) die("cannot create the worker pool");
		for (i = 0; i < n_threads; ++i)
			if (g_tbuf[i] == 0) g_tbuf[i] = mb_tbuf_init(1);
	}
	pthread_mutex_unlock(&g_mu);
	return g_pool;
}

/* ---- one mini-batch ---------------------------------------------------------------------------- */

typedef struct {
	const mb_opt_t *opt;
	const mb_idx_t *idx;
	int32_t n_seq;
	const int32_t *qlen;
	const char *const *seq;          /* ASCII (or 0-3) bases */
	const char *const *name;         /* may be null */
	const uint8_t *mt;               /* per read l2b_meth_t */
	int32_t *n_hit;
	mb_hit_t **hit;
	/* the front pass */
	uint8_t *nt4;
	int64_t *off;
	int32_t *len;
	uint8_t *no_chain;
	mbg_chain_opt_t *copt;
	int32_t *copt_idx, n_copt;
	mbx_front_t front;
	/* pairing */
	int32_t n_pair;
	const int32_t *pair_off;         /* pair k: reads pair_off[k], pair_off[k] + 1 */
	const mb_pestat_t *pes;
} batch_t;

/* per coroutine: the read being mapped and the anchors handed to its chaining call */
typedef struct {
	const batch_t *b;
	int64_t r;
	mb_anchor_t *anc_given;
	int64_t anc_n;
} tctx_t;
static __thread tctx_t *tls_tc;

static void post_and_wait(mbx_req_t *q)
{
	int tid = mbs_tid();
	tctx_t *tc = tls_tc;
	req_list_t *l = &g_posted[tid];
	if (l->n == l->m) {
		l->m = l->m ? l->m * 2 : 4096;
		l->a = (mbx_req_t**)realloc(l->a, l->m * sizeof(mbx_req_t*));
		if (l->a == 0) die("out of memory (request list)");
	}
	l->a[l->n++] = q;
	mbs_yield();
	tls_tc = tc;
}

/* one round: every posted request of every thread, as one batch per kind */
static mbx_req_t **g_round_buf;
static int64_t g_round_m;

static void round_cb(void *data, int n_threads)
{
	int64_t tot = 0, n_d2 = 0, n_z2 = 0, n_ll = 0, i;
	int t;
	(void)data;
	for (t = 0; t < n_threads; ++t) tot += g_posted[t].n;
	if (tot * 3 > g_round_m) {
		g_round_m = tot * 3 + 1024;
		g_round_buf = (mbx_req_t**)realloc(g_round_buf, g_round_m * sizeof(mbx_req_t*));
		if (g_round_buf == 0) die("out of memory (round)");
	}
	for (t = 0; t < n_threads; ++t) {
		req_list_t *l = &g_posted[t];
		for (i = 0; i < l->n; ++i) {
			mbx_req_t *q = l->a[i];
			if (q->kind == MBX_D2) g_round_buf[n_d2++] = q;
			else if (q->kind == MBX_Z2) g_round_buf[tot + n_z2++] = q;
			else g_round_buf[2 * tot + n_ll++] = q;
		}
		l->n = 0;
	}
	if (n_d2) mbx_run_dp(1, (int32_t)n_d2, g_round_buf);
	if (n_z2) mbx_run_dp(0, (int32_t)n_z2, g_round_buf + tot);
	if (n_ll) mbx_run_ll((int32_t)n_ll, g_round_buf + 2 * tot);
}

static void se_task(void *data, int64_t r, int tid)
{
	const batch_t *b = (const batch_t*)data;
	const mbx_front_t *f = &b->front;
	mb_tbuf_t *tb = g_tbuf[tid];
	tctx_t tc;
	mb_opt_t oa;
	mb_sai_v u;
	l2b_meth_t mt = (l2b_meth_t)b->mt[r];
	tc.b = b, tc.r = r, tc.anc_given = 0, tc.anc_n = 0;
	tls_tc = &tc;
	mb_opt_adap(b->opt, b->qlen[r], &oa);
	u.n = u.m = (size_t)f->n_intv[r];
	u.a = u.n ? (mb_sai_t*)(f->intv + f->intv_off[r]) : 0;
	if (u.n && f->n_anc[r] < 0) {      /* the reference's own mb_anchor sorts and frees u: give it a copy */
		mb_sai_t *c = (mb_sai_t*)(kmalloc)(mb_tbuf_km(tb), u.n * sizeof(mb_sai_t));   /* kalloc's: map-algo.c frees it */
		memcpy(c, u.a, u.n * sizeof(mb_sai_t));
		u.a = c;
	}
	b->hit[r] = mb_map_sai(&oa, b->idx, b->qlen[r], b->seq[r], mt, &u, &b->n_hit[r], tb, b->name ? b->name[r] : 0);
	tls_tc = 0;
}

static void pe_task(void *data, int64_t k, int tid)
{
	const batch_t *b = (const batch_t*)data;
	int32_t off = b->pair_off[k], len[2];
	char *seq[2];
	len[0] = b->qlen[off], len[1] = b->qlen[off + 1];
	seq[0] = (char*)b->seq[off], seq[1] = (char*)b->seq[off + 1];
	mb_pair(mb_tbuf_km(g_tbuf[tid]), b->opt, b->idx->l2b, &b->n_hit[off], &b->hit[off], b->pes, len, seq);
}

/* chaining options mb_map_sai passes to its first mb_lchain_dp call (map-algo.c:606-612) */
static void chain_opt_of(const mb_opt_t *opt, int32_t qlen, mbg_chain_opt_t *o)
{
	mb_opt_t oa;
	mb_opt_adap(opt, qlen, &oa);
	memset(o, 0, sizeof(*o));
	o->max_dist_x = o->max_dist_y = oa.max_gap;
	o->bw = oa.bw, o->max_skip = oa.max_chain_skip, o->max_iter = oa.max_chain_iter, o->min_sc = oa.min_chain_score;
	o->pen_gap = (float)(oa.chain_gap_scale * .01 * oa.min_len);
}

/* The front pass and the single-end mapping of a mini-batch. */
static void map_reads(batch_t *b, int n_threads)
{
	mbs_pool_t *pool = pool_for(n_threads);
	int64_t i, tot = 0;
	double t0 = now_s(), t1, t2;
	mbx_front_in_t in;
	int32_t last_len = -1;
	b->off = (int64_t*)malloc((b->n_seq + 1) * sizeof(int64_t));
	b->len = (int32_t*)malloc((b->n_seq + 1) * sizeof(int32_t));
	b->no_chain = (uint8_t*)calloc(b->n_seq + 1, 1);
	b->copt_idx = (int32_t*)malloc((b->n_seq + 1) * sizeof(int32_t));
	b->copt = 0, b->n_copt = 0;
	for (i = 0; i < b->n_seq; ++i) tot += b->qlen[i];
	b->nt4 = (uint8_t*)malloc(tot + 1);
	if (!b->off || !b->len || !b->no_chain || !b->copt_idx || !b->nt4) die("out of memory (mini-batch)");
	for (i = 0, tot = 0; i < b->n_seq; ++i) {
		int32_t l, n = b->qlen[i];
		const uint8_t *s = (const uint8_t*)b->seq[i];
		uint8_t *d = b->nt4 + tot;
		b->off[i] = tot, b->len[i] = n;
		for (l = 0; l < n; ++l) d[l] = kom_nt4_table[s[l]];
		if (b->mt[i] != L2B_METH_NONE) {
			l2b_meth_convert((l2b_meth_t)b->mt[i], n, d);
			b->no_chain[i] = 1;
		}
		tot += n;
		/* reads of one length share their chaining options; lengths mostly repeat in runs */
		if (n != last_len) {
			mbg_chain_opt_t o;
			int32_t k;
			chain_opt_of(b->opt, n, &o);
			for (k = 0; k < b->n_copt; ++k)
				if (memcmp(&b->copt[k], &o, sizeof(o)) == 0) break;
			if (k == b->n_copt) {
				if ((b->n_copt & (b->n_copt - 1)) == 0) {
					b->copt = (mbg_chain_opt_t*)realloc(b->copt, (b->n_copt ? b->n_copt * 2 : 1) * sizeof(o));
					if (b->copt == 0) die("out of memory (chain options)");
				}
				b->copt[b->n_copt++] = o;
			}
			b->copt_idx[i] = k, last_len = n;
		} else b->copt_idx[i] = b->copt_idx[i - 1];
	}
	b->off[b->n_seq] = tot;
	memset(&in, 0, sizeof(in));
	in.n = b->n_seq, in.nt4 = b->nt4, in.off = b->off, in.len = b->len, in.no_chain = b->no_chain;
	in.min_len = b->opt->min_len, in.max_sub_occ = b->opt->max_sub_occ, in.max_occ = b->opt->max_occ;
	in.n_copt = b->n_copt, in.copt = b->copt, in.copt_idx = b->copt_idx;
	mbx_front(&in, &b->front);
	t1 = now_s();
	mbs_run(pool, b->n_seq, 64, se_task, round_cb, b);
	t2 = now_s();
	if (g_stats) {
		mbs_stats_t st;
		mbs_get_stats(pool, &st);
		fprintf(stderr, "[mbg_map] %d reads: front %.3f s (device %.1f ms), SE %.3f s; rounds so far %ld, peak coroutines %ld\n",
		        b->n_seq, t1 - t0, b->front.dev_ms, t2 - t1, (long)st.rounds, (long)st.coroutines_peak);
	}
}

static void map_free(batch_t *b)
{
	free(b->off); free(b->len); free(b->no_chain); free(b->copt_idx); free(b->copt); free(b->nt4);
	b->off = 0, b->len = 0, b->no_chain = 0, b->copt_idx = 0, b->copt = 0, b->nt4 = 0;
}

/* Pairing of a mini-batch's pairs (b->pair_off, b->pes set). */
static void pair_reads(batch_t *b)
{
	double t0 = now_s();
	if (b->n_pair <= 0) return;
	mbs_run(g_pool, b->n_pair, 64, pe_task, round_cb, b);
	if (g_stats) fprintf(stderr, "[mbg_map] %d pairs paired in %.3f s\n", b->n_pair, now_s() - t0);
}

/* ---- the CLI: kt_pipeline's mapping step --------------------------------------------------------- */

void __real_kt_pipeline(int n_threads, void *(*func)(void*, int, void*), void *shared_data, int n_steps);

static struct {
	void *(*func)(void*, int, void*);
	void *shared;
} g_pipe;

static void *map_step(ref_step_t *s)
{
	const ref_pipeline_t *p = s->p;
	const mb_opt_t *opt = p->opt;
	const mb_idx_t *idx = p->idx;
	batch_t b;
	int32_t k, j, *pair_off = 0, n_pair = 0;
	uint8_t *mt;
	if (opt == 0 || idx == 0 || p->mb_size != opt->mb_size || p->n_fp < 1)
		die("the pipeline record does not match minibwa 0.7-r421's layout");
	mt = (uint8_t*)calloc(s->n_seq + 1, 1);
	for (k = 0; k < s->n_frag; ++k)       /* worker_for_se_batch's conversion (map-main.c:66) */
		for (j = 0; j < s->seg_cnt[k]; ++j)
			mt[s->seg_off[k] + j] = !idx->is_meth ? L2B_METH_NONE : (j & 1) == 0 ? L2B_METH_C2T : L2B_METH_G2A;
	memset(&b, 0, sizeof(b));
	b.opt = opt, b.idx = idx, b.n_seq = s->n_seq, b.mt = mt;
	{
		int32_t *ql = (int32_t*)malloc((s->n_seq + 1) * sizeof(int32_t));
		const char **sq = (const char**)malloc((s->n_seq + 1) * sizeof(char*));
		const char **nm = (const char**)malloc((s->n_seq + 1) * sizeof(char*));
		for (k = 0; k < s->n_seq; ++k)
			ql[k] = (int32_t)s->seq[k].l_seq, sq[k] = s->seq[k].seq, nm[k] = s->seq[k].name;
		b.qlen = ql, b.seq = sq, b.name = nm;
	}
	b.n_hit = s->n_hit, b.hit = s->hit;
	map_reads(&b, opt->n_thread);
	if ((opt->flag & MB_F_PE) && s->n_frag < s->n_seq && !(opt->flag & (MB_F_NO_PAIRING | MB_F_NO_ALN))) { /* map-main.c:180-192 */
		if ((opt->flag & MB_F_PE_PREDEF) || s->n_pe < 20) {
			s->pes[1].failed = 0;
			s->pes[1].avg = opt->pe_avg, s->pes[1].std = opt->pe_std;
			s->pes[1].lo = opt->pe_lo, s->pes[1].hi = opt->pe_hi;
		} else {
			void *km = opt->flag & MB_F_NO_KALLOC ? 0 : km_init();
			mb_pestat(km, opt, s->n_frag, s->seg_off, s->seg_cnt, s->n_hit, s->hit, s->pes);
			if (km) km_destroy(km);
		}
		pair_off = (int32_t*)malloc((s->n_frag + 1) * sizeof(int32_t));
		for (k = 0; k < s->n_frag; ++k)
			if (s->seg_cnt[k] == 2) pair_off[n_pair++] = s->seg_off[k];
		b.n_pair = n_pair, b.pair_off = pair_off, b.pes = s->pes;
		pair_reads(&b);
	}
	map_free(&b);
	free(pair_off);
	free((void*)b.qlen); free((void*)b.seq); free((void*)b.name); free(mt);
	return s;
}

static void *pipe_step(void *data, int step, void *in)
{
	(void)data;
	if (step == 1) return map_step((ref_step_t*)in);
	return g_pipe.func(g_pipe.shared, step, in);
}

void __wrap_kt_pipeline(int n_threads, void *(*func)(void*, int, void*), void *shared_data, int n_steps)
{
	if (n_steps != 3) {          /* not the mapper's three-step pipeline */
		__real_kt_pipeline(n_threads, func, shared_data, n_steps);
		return;
	}
	g_pipe.func = func, g_pipe.shared = shared_data;
	__real_kt_pipeline(n_threads, pipe_step, 0, n_steps);
}

/* ---- the public batch API (map-algo.c:703-770) ---------------------------------------------------- */

mb_hit_t **__wrap_mb_map_batch(const mb_opt_t *opt, const mb_idx_t *idx, int32_t n_seq, const int32_t *qlen, const char **seq,
                               int32_t *n_hit, mb_tbuf_t *b0, const char **qname)
{
	int32_t i, is_pe = !!(opt->flag & MB_F_PE), is_meth = !!(opt->flag & MB_F_METH), *pair_off = 0, n_pair = 0;
	mb_hit_t **hit;
	uint8_t *mt;
	batch_t b;
	mb_pestat_t pes[4];
	const char *env = getenv("MBG_MAP_THREADS");
	(void)b0;
	if (n_seq <= 0) return 0;
	if (is_meth && !idx->is_meth) return 0;
	hit = (mb_hit_t**)calloc(n_seq, sizeof(mb_hit_t*));
	mt = (uint8_t*)calloc(n_seq, 1);
	for (i = 0; i < n_seq; ++i)
		mt[i] = !is_meth ? L2B_METH_NONE : !is_pe || (i & 1) == 0 ? L2B_METH_C2T : L2B_METH_G2A;
	memset(&b, 0, sizeof(b));
	b.opt = opt, b.idx = idx, b.n_seq = n_seq, b.qlen = qlen, b.seq = seq, b.name = qname, b.mt = mt;
	b.n_hit = n_hit, b.hit = hit;
	for (i = 0; i < n_seq; ++i) n_hit[i] = 0;
	map_reads(&b, env ? atoi(env) : opt->n_thread);
	if (is_pe && n_seq >= 2 && !(opt->flag & (MB_F_NO_PAIRING | MB_F_NO_ALN))) {
		for (i = 0; i < 4; ++i) pes[i].failed = 1;
		pes[1].failed = 0;
		pes[1].avg = opt->pe_avg, pes[1].std = opt->pe_std;
		pes[1].lo = opt->pe_lo, pes[1].hi = opt->pe_hi;
		pair_off = (int32_t*)malloc((n_seq / 2 + 1) * sizeof(int32_t));
		for (i = 0; i + 1 < n_seq; i += 2) pair_off[n_pair++] = i;
		b.n_pair = n_pair, b.pair_off = pair_off, b.pes = pes;
		pair_reads(&b);
	}
	map_free(&b);
	free(pair_off);
	free(mt);
	return hit;
}

/* ---- index --------------------------------------------------------------------------------------- */

mb_idx_t *__real_mb_idx_load(const char *prefix, int32_t is_meth);
mb_idx_t *__wrap_mb_idx_load(const char *prefix, int32_t is_meth)
{
	mb_idx_t *idx = __real_mb_idx_load(prefix, is_meth);
	if (idx) {
		pthread_mutex_lock(&g_mu);
		mbx_index(idx);
		g_idx = idx;
		pthread_mutex_unlock(&g_mu);
	}
	return idx;
}

static void need_index(const mb_bwt_t *bwt)
{
	if (g_idx == 0 || (bwt && bwt != g_idx->bwt)) die("the index was not loaded through mb_idx_load");
}

/* ---- anchors and chains from the front pass ------------------------------------------------------ */

double __real_mb_anchor(void *km, const mb_idx_t *idx, mb_sai_v *u, int32_t min_len, int32_t qlen, const uint8_t *qseq,
                        l2b_meth_t mt, int32_t max_occ, mb_anchor_v *v);
double __wrap_mb_anchor(void *km, const mb_idx_t *idx, mb_sai_v *u, int32_t min_len, int32_t qlen, const uint8_t *qseq,
                        l2b_meth_t mt, int32_t max_occ, mb_anchor_v *v)
{
	tctx_t *tc = tls_tc;
	if (tc && mt == L2B_METH_NONE && idx == tc->b->idx && max_occ == tc->b->opt->max_occ && tc->b->front.n_anc[tc->r] >= 0 &&
	    (u->n == 0 || u->a == tc->b->front.intv + tc->b->front.intv_off[tc->r])) {
		const mbx_front_t *f = &tc->b->front;
		int64_t n = f->n_anc[tc->r];
		v->n = 0;
		if (u->n == 0) return 1.0;
		if (n > v->m) {
			v->m = n;
			v->a = (mb_anchor_t*)(krealloc)(km, v->a, (size_t)v->m * sizeof(mb_anchor_t));
		}
		if (n) memcpy(v->a, f->anc + f->anc_off[tc->r], (size_t)n * sizeof(mb_anchor_t));
		v->n = n;
		u->a = 0, u->n = u->m = 0;       /* the shared intervals are not the caller's to free (map-algo.c:602) */
		tc->anc_given = v->a, tc->anc_n = n;
		return 1.0;                      /* seed_ratio of an unconverted read (seed.c:276) */
	}
	return __real_mb_anchor(km, idx, u, min_len, qlen, qseq, mt, max_occ, v);
}

mb_anchor_t *__real_mb_lchain_dp(void *km, const l2b_t *l2b, int max_dist_x, int max_dist_y, int bw, int max_skip, int max_iter,
                                 int min_sc, float chn_pen_gap, int64_t n, mb_anchor_t *a, int *n_u_, uint64_t **_u);
mb_anchor_t *__wrap_mb_lchain_dp(void *km, const l2b_t *l2b, int max_dist_x, int max_dist_y, int bw, int max_skip, int max_iter,
                                 int min_sc, float chn_pen_gap, int64_t n, mb_anchor_t *a, int *n_u_, uint64_t **_u)
{
	tctx_t *tc = tls_tc;
	if (tc && tc->anc_given && a == tc->anc_given && n == tc->anc_n) {
		const batch_t *b = tc->b;
		const mbg_chain_opt_t *o = &b->copt[b->copt_idx[tc->r]];
		if (o->max_dist_x == max_dist_x && o->max_dist_y == max_dist_y && o->bw == bw && o->max_skip == max_skip &&
		    o->max_iter == max_iter && o->min_sc == min_sc && o->pen_gap == chn_pen_gap) {
			const mbx_front_t *f = &b->front;
			int32_t n_u = f->n_u[tc->r];
			int64_t n_c = f->n_chn[tc->r];
			uint64_t *u;
			mb_anchor_t *c;
			tc->anc_given = 0;
			(kfree)(km, a);                /* the input is consumed (lchain.c:139) */
			if (_u) *_u = 0, *n_u_ = 0;
			if (n_u == 0) return 0;
			u = (uint64_t*)(kmalloc)(km, (size_t)n_u * sizeof(uint64_t));
			c = (mb_anchor_t*)(kmalloc)(km, (size_t)n_c * sizeof(mb_anchor_t));
			memcpy(u, f->u + f->anc_off[tc->r], (size_t)n_u * sizeof(uint64_t));
			memcpy(c, f->chn + f->anc_off[tc->r], (size_t)n_c * sizeof(mb_anchor_t));
			if (_u) *_u = u, *n_u_ = n_u;
			else (kfree)(km, u);
			return c;
		}
	}
	if (tc) tc->anc_given = 0;
	return __real_mb_lchain_dp(km, l2b, max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc, chn_pen_gap, n, a, n_u_, _u);
}

/* ---- deferred DP and local SW -------------------------------------------------------------------- */

static pthread_mutex_t g_solo_mu = PTHREAD_MUTEX_INITIALIZER;

static void dp_call(int dual, void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                    int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez)
{
	mbx_req_t r;
	int solo = !mbs_in_task();
	memset(&r, 0, sizeof(r));
	if (m < 1 || m > 5) die("DP with an alphabet larger than 5");
	r.kind = dual ? MBX_D2 : MBX_Z2;
	r.qlen = qlen > 0 ? qlen : 0, r.tlen = tlen > 0 ? tlen : 0;   /* the reference returns the reset state */
	r.q = query, r.t = target, r.m = m;
	memcpy(r.mat, mat, (size_t)m * m);
	r.gq = q, r.ge = e, r.gq2 = q2, r.ge2 = e2;
	r.w = w, r.zdrop = zdrop, r.end_bonus = end_bonus, r.flag = flag;
	if (solo) {
		mbx_req_t *p = &r;
		pthread_mutex_lock(&g_solo_mu);
		mbx_run_dp(dual, 1, &p);
	} else post_and_wait(&r);
	ez->max = (uint32_t)r.dp.max & 0x7fffffff, ez->zdropped = r.dp.zdropped ? 1 : 0;
	ez->max_q = r.dp.max_q, ez->max_t = r.dp.max_t;
	ez->mqe = r.dp.mqe, ez->mqe_t = r.dp.mqe_t;
	ez->mte = r.dp.mte, ez->mte_q = r.dp.mte_q;
	ez->score = r.dp.score, ez->reach_end = r.dp.reach_end;
	ez->n_cigar = r.dp.n_cigar;
	if (r.dp.n_cigar > ez->m_cigar) {
		ez->m_cigar = r.dp.n_cigar;
		ez->cigar = (uint32_t*)krealloc(km, ez->cigar, (size_t)ez->m_cigar * 4);   /* ksw2.h's allocator, as ksw2 itself */
	}
	if (r.dp.n_cigar > 0) memcpy(ez->cigar, r.cigar, (size_t)r.dp.n_cigar * 4);
	if (solo) pthread_mutex_unlock(&g_solo_mu);
}

void __wrap_ksw_extd2_sse(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                          int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez)
{
	dp_call(1, km, qlen, query, tlen, target, m, mat, q, e, q2, e2, w, zdrop, end_bonus, flag, ez);
}

void __wrap_ksw_extz2_sse(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                          int8_t q, int8_t e, int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez)
{
	dp_call(0, km, qlen, query, tlen, target, m, mat, q, e, q, e, w, zdrop, end_bonus, flag, ez);
}

/* ksw_ll_qinit's query profile is kept as the query and matrix themselves; the core call runs the job.
 * The record comes from the caller's km with ksw2.h's allocator, as the caller's kfree (pe.c:281,
 * align.c:797) expects. */
typedef struct { int32_t qlen, m, size; int8_t mat[25]; uint8_t q[]; } ll_query_t;

void *__wrap_ksw_ll_qinit(void *km, int size, int qlen, const uint8_t *query, int m, const int8_t *mat)
{
	ll_query_t *q = (ll_query_t*)kmalloc(km, sizeof(ll_query_t) + (qlen > 0 ? qlen : 0) + 1);
	if (m < 1 || m > 5) die("ksw_ll_qinit with an alphabet larger than 5");
	q->qlen = qlen > 0 ? qlen : 0, q->m = m, q->size = size;
	memcpy(q->mat, mat, (size_t)m * m);
	if (q->qlen) memcpy(q->q, query, q->qlen);
	return q;
}

static ksw_llrst_t ll_call(void *q_, int tlen, const uint8_t *target, int gapo, int gape, int xtra, int size)
{
	const ll_query_t *q = (const ll_query_t*)q_;
	mbx_req_t r;
	ksw_llrst_t o;
	memset(&r, 0, sizeof(r));
	r.kind = MBX_LL;
	r.qlen = q->qlen, r.tlen = tlen > 0 ? tlen : 0;
	r.q = q->q, r.t = target, r.m = (int8_t)q->m;
	memcpy(r.mat, q->mat, (size_t)q->m * q->m);
	r.size = size, r.gapo = gapo, r.gape = gape, r.xtra = xtra;
	if (!mbs_in_task()) {
		mbx_req_t *p = &r;
		pthread_mutex_lock(&g_solo_mu);
		mbx_run_ll(1, &p);
		pthread_mutex_unlock(&g_solo_mu);
	} else post_and_wait(&r);
	o.score = r.ll.score, o.te = r.ll.te, o.qe = r.ll.qe, o.score2 = r.ll.score2, o.te2 = r.ll.te2;
	return o;
}

ksw_llrst_t __wrap_ksw_ll_u8_core(void *q, int tlen, const uint8_t *target, int gapo, int gape, int xtra)
{
	return ll_call(q, tlen, target, gapo, gape, xtra, 1);
}

ksw_llrst_t __wrap_ksw_ll_i16_core(void *q, int tlen, const uint8_t *target, int gapo, int gape, int xtra)
{
	return ll_call(q, tlen, target, gapo, gape, xtra, 2);
}

/* ksw_ll_i16 (ksw2_ll_sse.c:335-341) calls the i16 core inside its own file, so it is bound here too */
int __wrap_ksw_ll_i16(void *q, int tlen, const uint8_t *target, int gapo, int gape, int *qe, int *te)
{
	ksw_llrst_t r = ll_call(q, tlen, target, gapo, gape, 0, 2);
	*qe = r.qe, *te = r.te;
	return r.score;
}

/* ---- hot-path calls outside the batched driver --------------------------------------------------- */

void __wrap_mb_bwt_sa_batch(void *km, const mb_bwt_t *bwt, int64_t n, uint64_t *x)
{
	(void)km;
	if (n <= 0) return;
	need_index(bwt);
	pthread_mutex_lock(&g_solo_mu);
	mbx_sa(n, x);
	pthread_mutex_unlock(&g_solo_mu);
}

void __wrap_mb_seed_intv_batch(void *km, const mb_bwt_t *bwt, int32_t n_seq, const int32_t *len, uint8_t *const *seq, int32_t min_len,
                               int32_t max_sub_occ, mb_sai_v *v)
{
	if (n_seq <= 0) return;
	need_index(bwt);
	pthread_mutex_lock(&g_solo_mu);
	mbx_seed(km, n_seq, len, seq, min_len, max_sub_occ, v);
	pthread_mutex_unlock(&g_solo_mu);
}

void __wrap_mb_seed_intv(void *km, const mb_bwt_t *bwt, int32_t len, const uint8_t *seq, int32_t min_len, int32_t max_sub_occ, mb_sai_v *v)
{
	uint8_t *s = (uint8_t*)seq;
	v->n = 0;
	if (len <= 0) return;
	need_index(bwt);
	pthread_mutex_lock(&g_solo_mu);
	mbx_seed(km, 1, &len, &s, min_len, max_sub_occ, v);
	pthread_mutex_unlock(&g_solo_mu);
}
