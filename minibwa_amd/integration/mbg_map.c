/*
 * mbg_map.c -- the batched drop-in mapper: the reference CLI (`minibwa map`, `minibwa mem`) and its
 * public batch API (mb_map_batch, minibwa.h:170) with one device pass per mini-batch for seeding,
 * anchors and chains, and every DP and mate-rescue local-SW call of the mini-batch executed in device
 * rounds.  The reference's own host code -- option parsing, reading, mb_map_sai's hit bookkeeping,
 * mb_align1, pairing, rescue and SAM/PAF formatting -- runs unchanged; this file replaces only the
 * mapping step of its pipeline and binds the hook points below.
 *
 * Linked into the unmodified reference objects with -Wl,--wrap for:
 *   kt_pipeline          (kthread.h:10; only caller mb_map_file, map-main.c:292)  worker_pipeline's three
 *                         steps (map-main.c:112-256) run as read_step, map_step and write_step below,
 *                         still overlapped by the reference's kt_pipeline
 *   mb_bseq_open         (bseq.h:20)  records each input's name so that read_step can read the same
 *                         files with the native reader (mbg_fq_read / mbg_fq_read_frag, include/mbg.h),
 *                         whose records equal mb_bseq_read / mb_bseq_read_frag's
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
 *              yields; the scheduler's dispatcher runs the posted jobs as device batches while the
 *              threads go on with other reads, and each read resumes with its result -- so a read's
 *              calls keep their order and their dependences (the gap-fill loop's z-drop break and
 *              second pass, align.c:689-721, see the previous call's result);
 *   3. PE      the insert-size model (mb_pestat, or the predefined one) exactly as worker_pipeline
 *              chooses it (map-main.c:180-190), then mb_pair per pair in coroutines: each rescue
 *              candidate's ksw_ll and extension DP are deferred into rounds the same way.
 * Results are the reference's by construction: the same functions run on the same data, and every
 * deferred call returns what the reference's call returns (the device kernels are bit-exact).
 *
 * Mini-batches are read, mapped and written as worker_pipeline does: the same chunking (mb_bseq_read(_frag)
 * rules), the same fragments and name trimming, the same insert-size model per mini-batch, the same
 * records in the same order.  Formatting (mb_format's records, map-main.c:207-239, written by
 * mbg_fmt.c) runs on a second pool in fragment ranges whose strings the output step writes in order.
 *
 * Groups (mbg_map.h) are the serving loop's unit: a set of reads whose front pass has run, mapped one
 * read or one pair per task inside an open run of the worker pool, pairing included (predefined
 * insert-size statistics), so that reads of several groups map at the same time.
 *
 * The layout of pipeline_t (map-main.c:16-24, minibwa 0.7-r421) is restated below because kt_pipeline
 * hands it over as void*; the pipeline checks that its fields agree with its options before use.
 *
 * Environment: MBG_MAP_THREADS (pool size for mb_map_batch; the CLI uses -t), MBG_MAP_STATS=1 (per
 * mini-batch timing on stderr).
 */
#define _GNU_SOURCE
#include <assert.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "kalloc.h"
#include "bwt.h"
#include "bseq.h"
#include "mbpriv.h"
#include "kthread.h"
#include "ksw2.h"
#include "mbg.h"
#include "mbg_exec.h"
#include "mbg_fmt.h"
#include "mbg_pe.h"
#include "mbg_map.h"
#include "mbg_sched.h"

/* ---- the reference's pipeline record (map-main.c:16-24) ----------------------------------------- */

typedef struct {
	int32_t n_fp, n_threads;
	int64_t n_base, n_seq;
	int64_t mb_size;
	const mb_opt_t *opt;
	mb_bseq_file_t **fp;
	const mb_idx_t *idx;
	FILE *fp_out;
} ref_pipeline_t;


/* ---- global state ------------------------------------------------------------------------------ */

#define MAX_THREADS 256

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;      /* pool creation, calls outside rounds */
static mbs_pool_t *g_pool;
static mb_tbuf_t *g_tbuf[MAX_THREADS];
static const mb_idx_t *g_idx;
static int g_stats;
static int64_t g_flush = 4096;    /* pending requests that start a device batch (MBG_FLUSH) */
static int g_fmt_native;          /* output records by mbg_fmt_record (else the reference's mb_format) */
static int g_fmt_check;           /* MBG_FMT_CHECK=1: every record by both, compared */

static double now_s(void);

/* MBG_TIMELINE=1: the start and end of every pipeline stage of every mini-batch, printed after the
 * pipeline, relative to its start */
#define TL_CAP 4096
static int g_tl;
static double g_tl_t0;
static atomic_int g_tl_n;
static struct { double t; int64_t batch; const char *ev; } g_tl_ev[TL_CAP];

static void tl_mark(int64_t batch, const char *ev)
{
	int k;
	if (!g_tl) return;
	k = atomic_fetch_add(&g_tl_n, 1);
	if (k < TL_CAP) g_tl_ev[k].t = now_s(), g_tl_ev[k].batch = batch, g_tl_ev[k].ev = ev;
}

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
	const mbg_counters_t *d = &xs.dev;
	memset(&ss, 0, sizeof(ss));
	if (g_pool) mbs_get_stats(g_pool, &ss);
	mbx_get_stats(&xs);
	fprintf(stderr, "[mbg_map] summary: executor %s formatter %s fronts %ld reads %ld front_s %.3f | batches %ld waits %ld peak_coroutines %ld "
	        "batch_s %.3f run_s %.3f busy_s %.3f idle_s %.3f | dp_calls %ld dp_jobs %ld dp_s %.3f dp_pack_s %.3f | ll_calls %ld "
	        "ll_jobs %ld ll_s %.3f | "
	        "dev_seed1_ms %.3f dev_seed2_ms %.3f dev_group_ms %.3f dev_anchor_ms %.3f dev_chain_ms %.3f dev_front_copy_ms %.3f "
	        "dev_front_ms %.3f seed_records %ld anchor_slots %ld seed1_occ_blocks %lu seed1_kmer_lookups %lu dev_dp_launches %ld "
	        "dev_dp_jobs %ld dev_dp_cells %ld dev_dp_ms %.3f dev_ll_launches %ld dev_ll_cells %ld dev_ll_ms %.3f graph_launches %ld "
	        "graph_instantiations %ld direct_chains %ld\n",
	        mbx_name(), g_fmt_native ? "native" : "reference", (long)xs.fronts, (long)xs.front_reads, xs.front_s, (long)ss.batches, (long)ss.waits, (long)ss.coroutines_peak,
	        ss.batch_s, ss.task_s, ss.busy_s, ss.idle_s, (long)xs.dp_calls, (long)xs.dp_jobs, xs.dp_s, xs.dp_pack_s, (long)xs.ll_calls,
	        (long)xs.ll_jobs, xs.ll_s,
	        d->ms_seed1, d->ms_seed2, d->ms_group, d->ms_anchor, d->ms_chain, d->ms_front_copy, d->ms_front, (long)d->front_records,
	        (long)d->front_slots, (unsigned long)d->seed1_occ_blocks, (unsigned long)d->seed1_kmer_lookups, (long)d->dp_launches,
	        (long)d->dp_jobs, (long)d->dp_cells, d->ms_dp, (long)d->ll_launches, (long)d->ll_cells, d->ms_ll, (long)d->graph_launches,
	        (long)d->graph_instantiations, (long)d->direct_chains);
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
		s = getenv("MBG_FLUSH");
		if (s && atol(s) > 0) g_flush = atol(s);
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
	double t_front;
	int64_t seq_no;               /* the mini-batch's number in the run (for the timeline) */
	int front_done;               /* the front pass has run (set by the front thread, under g_front.mu) */
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
	tctx_t *tc = tls_tc;
	mbs_wait(q);
	tls_tc = tc;
}

/* one device batch: every request the dispatcher took, one library call per kind */
static mbx_req_t **g_round_buf;
static int64_t g_round_m;

static void round_cb(void *data, void **reqs, int64_t tot)
{
	int64_t n_d2 = 0, n_z2 = 0, n_ll = 0, i;
	(void)data;
	if (tot * 3 > g_round_m) {
		g_round_m = tot * 3 + 1024;
		g_round_buf = (mbx_req_t**)realloc(g_round_buf, g_round_m * sizeof(mbx_req_t*));
		if (g_round_buf == 0) die("out of memory (round)");
	}
	for (i = 0; i < tot; ++i) {
		mbx_req_t *q = (mbx_req_t*)reqs[i];
		if (q->kind == MBX_D2) g_round_buf[n_d2++] = q;
		else if (q->kind == MBX_Z2) g_round_buf[tot + n_z2++] = q;
		else g_round_buf[2 * tot + n_ll++] = q;
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

/* The front pass of a mini-batch on front context `slot`.  nt4/off (may be null): the reads' nt4 codes
 * as the native reader left them (kom_nt4_table's), used as they are for unconverted reads. */
static void front_reads(batch_t *b, int slot, const uint8_t *nt4, const int64_t *off)
{
	int64_t i, tot = 0;
	double t0 = now_s();
	mbx_front_in_t in;
	tl_mark(b->seq_no, "front+");
	int32_t last_len = -1;
	int any_meth = 0;
	b->off = (int64_t*)malloc((b->n_seq + 1) * sizeof(int64_t));
	b->len = (int32_t*)malloc((b->n_seq + 1) * sizeof(int32_t));
	b->no_chain = (uint8_t*)calloc(b->n_seq + 1, 1);
	b->copt_idx = (int32_t*)malloc((b->n_seq + 1) * sizeof(int32_t));
	b->copt = 0, b->n_copt = 0, b->nt4 = 0;
	if (!b->off || !b->len || !b->no_chain || !b->copt_idx) die("out of memory (mini-batch)");
	for (i = 0; i < b->n_seq; ++i) {
		tot += b->qlen[i];
		if (b->mt[i] != L2B_METH_NONE) any_meth = 1;
	}
	if (nt4 == 0 || any_meth) {
		b->nt4 = (uint8_t*)malloc(tot + 1);
		if (b->nt4 == 0) die("out of memory (mini-batch)");
	}
	for (i = 0, tot = 0; i < b->n_seq; ++i) {
		int32_t n = b->qlen[i];
		if (b->nt4) {
			uint8_t *d = b->nt4 + tot;
			if (nt4) memcpy(d, nt4 + off[i], n);
			else {
				const uint8_t *q = (const uint8_t*)b->seq[i];
				int32_t l;
				for (l = 0; l < n; ++l) d[l] = kom_nt4_table[q[l]];
			}
			if (b->mt[i] != L2B_METH_NONE) {
				l2b_meth_convert((l2b_meth_t)b->mt[i], n, d);
				b->no_chain[i] = 1;
			}
			b->off[i] = tot;
		} else b->off[i] = off[i];
		b->len[i] = n;
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
	b->off[b->n_seq] = b->nt4 ? tot : (b->n_seq ? off[b->n_seq - 1] + b->qlen[b->n_seq - 1] : 0);
	memset(&in, 0, sizeof(in));
	in.slot = slot;
	in.n = b->n_seq, in.nt4 = b->nt4 ? b->nt4 : nt4, in.off = b->off, in.len = b->len, in.no_chain = b->no_chain;
	in.min_len = b->opt->min_len, in.max_sub_occ = b->opt->max_sub_occ, in.max_occ = b->opt->max_occ;
	in.n_copt = b->n_copt, in.copt = b->copt, in.copt_idx = b->copt_idx;
	mbx_front(&in, &b->front);
	b->t_front = now_s() - t0;
	tl_mark(b->seq_no, "front-");
}

/* The single-end mapping of a mini-batch whose front pass has run. */
static void se_reads(batch_t *b, int n_threads)
{
	mbs_pool_t *pool = pool_for(n_threads);
	double t0 = now_s(), t1;
	tl_mark(b->seq_no, "se+");
	mbs_run(pool, b->n_seq, 64, se_task, round_cb, b, g_flush);
	t1 = now_s();
	tl_mark(b->seq_no, "se-");
	if (g_stats) {
		mbs_stats_t st;
		mbs_get_stats(pool, &st);
		fprintf(stderr, "[mbg_map] %d reads: front %.3f s (device %.1f ms), SE %.3f s; batches so far %ld, peak coroutines %ld, "
		        "thread busy %.3f s, idle %.3f s, batch callbacks %.3f s (cumulative)\n",
		        b->n_seq, b->t_front, b->front.dev_ms, t1 - t0, (long)st.batches, (long)st.coroutines_peak, st.busy_s, st.idle_s,
		        st.batch_s);
	}
}

static void map_reads(batch_t *b, int n_threads)
{
	front_reads(b, 0, 0, 0);
	se_reads(b, n_threads);
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
	tl_mark(b->seq_no, "pair+");
	mbs_run(g_pool, b->n_pair, 64, pe_task, round_cb, b, g_flush);
	tl_mark(b->seq_no, "pair-");
	if (g_stats) fprintf(stderr, "[mbg_map] %d pairs paired in %.3f s\n", b->n_pair, now_s() - t0);
}

/* ---- the CLI: kt_pipeline's three steps ---------------------------------------------------------- */

void __real_kt_pipeline(int n_threads, void *(*func)(void*, int, void*), void *shared_data, int n_steps);
mb_bseq_file_t *__real_mb_bseq_open(const char *fn);

/* the file name behind each reader mb_open_bseqs opened (map-main.c:257-272), so that the mini-batches
 * can be read by the native reader instead */
#define MAX_INPUTS 64
static struct { mb_bseq_file_t *fp; char *fn; } g_inputs[MAX_INPUTS];
static int g_n_inputs;

mb_bseq_file_t *__wrap_mb_bseq_open(const char *fn)
{
	mb_bseq_file_t *fp = __real_mb_bseq_open(fn);
	if (fp) {
		pthread_mutex_lock(&g_mu);
		g_inputs[g_n_inputs % MAX_INPUTS].fp = fp;
		free(g_inputs[g_n_inputs % MAX_INPUTS].fn);
		g_inputs[g_n_inputs % MAX_INPUTS].fn = strdup(fn ? fn : "-");
		++g_n_inputs;
		pthread_mutex_unlock(&g_mu);
	}
	return fp;
}

static struct {
	const ref_pipeline_t *p;
	mbg_fq_t *fq[MAX_INPUTS];
	int n_fq;
	int64_t n_seq;                /* reads so far: each read's id (map-main.c:119) */
	int64_t n_batch;              /* mini-batches read so far: the front slot of the next one */
} g_pipe;
static mbs_pool_t *g_fmt_pool;    /* formatting in the output step, beside the mapping pool */

/* the record formatter needs the run's read group, which only format.c knows: learnt from a probe record */
void mbg_map_fmt_init(const mb_opt_t *opt)
{
	const char *e = getenv("MBG_FMT");
	g_fmt_native = !(e && strcmp(e, "ref") == 0) && mbg_fmt_learn_rg(opt) == 0;
	e = getenv("MBG_FMT_CHECK");
	g_fmt_check = e && atoi(e) > 0;
}

/* one record into out, as mb_format (format.c:355-361) writes it */
static void fmt_one(void *km, kstring_t *out, const l2b_t *l2b, const mb_bseq1_t *t, int32_t n_seg, const int32_t *n_hit,
                    mb_hit_t *const *hit, int32_t hit_idx, const mb_opt_t *opt, int seg_idx, int32_t mate_qlen)
{
	size_t l0 = out->l;
	if (!g_fmt_native || mbg_fmt_record(out, l2b, t, n_seg, n_hit, hit, hit_idx, opt, seg_idx, mate_qlen) < 0) {
		mb_format(km, out, l2b, t, n_seg, n_hit, hit, hit_idx, opt, seg_idx, mate_qlen);
		return;
	}
	if (g_fmt_check) {
		kstring_t r = {0, 0, 0};
		mb_format(km, &r, l2b, t, n_seg, n_hit, hit, hit_idx, opt, seg_idx, mate_qlen);
		if (r.l != out->l - l0 || memcmp(r.s, out->s + l0, r.l) != 0) {
			fprintf(stderr, "[mbg_map] record of '%s' differs from mb_format:\n  ours: %.*s  ref:  %s", t->name, (int)(out->l - l0), out->s + l0, r.s);
			abort();
		}
		free(r.s);
	}
}

typedef struct {
	mbg_fq_buf_t *buf;
	mbg_fq_chunk_t ch;
	int32_t n_seq, n_frag, n_pe;
	mb_bseq1_t *seq;              /* names and comments in arena (NUL-terminated); bases and qualities in the chunk */
	char *arena;
	int32_t *seg_off, *seg_cnt, *n_hit;
	mb_hit_t **hit;
	mb_pestat_t pes[4];
	kstring_t *out;               /* the formatted records, in read order, one string per format task */
	int32_t n_out;
	int64_t tot_len;
	int64_t id0;                  /* the id of the chunk's first read */
	uint8_t *same;                /* same[i]: read i joins read i-1's fragment (the read step only) */
	double t_read, t_map, t_fmt;
	/* the mapping of the mini-batch (its front pass runs in the read step) */
	batch_t b;
	int32_t *qlen;
	const char **sq, **nm;
	uint8_t *mt;
} pbatch_t;

static void die_rc(const char *what, int rc)
{
	fprintf(stderr, "[mbg_map] %s failed (mbg error %d)\n", what, rc);
	abort();
}

/* The CLI's front passes run on threads of their own: the read step queues its mini-batch and
 * returns, so that reading the next mini-batches, their front passes (mostly device time) and the
 * previous one's mapping overlap; the map step waits for its batch's front.  Fronts are taken in
 * mini-batch order by MBG_FRONT_THREADS threads (one fewer than the mini-batches in flight by
 * default), each front on its own front slot and context (a slot is reused only after its previous
 * batch has been written: the pipeline holds at most as many mini-batches as there are slots), so that
 * fronts share the device: a long-read front is bound by the latency of its longest dependent SMEM
 * walks, not by the device's throughput, and the next fronts' walks fill the time.
 * MBG_FRONT_THREADS=0 (or MBG_FRONT_THREAD=0) runs the front inside the read step instead. */
#define FRONT_Q 8
#define FRONT_MAX_THREADS 5
/* mini-batches in flight in the CLI's pipeline (MBG_PIPE_DEPTH; the reference keeps three, kt_pipeline's
 * worker count, map-main.c:290-292): each holds a front slot, so at most MBX_FRONT_SLOTS.  The order of
 * the output and each mini-batch's contents do not depend on it. */
static int g_pipe_depth = 3;
static struct {
	pthread_t th[FRONT_MAX_THREADS];
	int n_th;
	pthread_mutex_t mu;
	pthread_cond_t cv_q, cv_done;
	int started, stop, n, head;
	struct { batch_t *b; int slot; const uint8_t *nt4; const int64_t *off; } q[FRONT_Q];
} g_front = {.mu = PTHREAD_MUTEX_INITIALIZER, .cv_q = PTHREAD_COND_INITIALIZER, .cv_done = PTHREAD_COND_INITIALIZER};

static void *front_main(void *arg)
{
	(void)arg;
	pthread_setname_np(pthread_self(), "mbg-front");
	for (;;) {
		batch_t *b;
		int slot;
		const uint8_t *nt4;
		const int64_t *off;
		pthread_mutex_lock(&g_front.mu);
		while (g_front.n == 0 && !g_front.stop) pthread_cond_wait(&g_front.cv_q, &g_front.mu);
		if (g_front.n == 0) {
			pthread_mutex_unlock(&g_front.mu);
			break;
		}
		b = g_front.q[g_front.head].b, slot = g_front.q[g_front.head].slot;
		nt4 = g_front.q[g_front.head].nt4, off = g_front.q[g_front.head].off;
		g_front.head = (g_front.head + 1) % FRONT_Q, --g_front.n;
		pthread_mutex_unlock(&g_front.mu);
		front_reads(b, slot, nt4, off);
		pthread_mutex_lock(&g_front.mu);
		b->front_done = 1;
		pthread_cond_broadcast(&g_front.cv_done);
		pthread_mutex_unlock(&g_front.mu);
	}
	return 0;
}

static void front_queue(batch_t *b, int slot, const uint8_t *nt4, const int64_t *off)
{
	pthread_mutex_lock(&g_front.mu);
	if (g_front.n == FRONT_Q) die("front queue full");
	{
		const int k = (g_front.head + g_front.n) % FRONT_Q;
		g_front.q[k].b = b, g_front.q[k].slot = slot, g_front.q[k].nt4 = nt4, g_front.q[k].off = off;
	}
	++g_front.n;
	pthread_cond_signal(&g_front.cv_q);
	pthread_mutex_unlock(&g_front.mu);
}

static void front_wait(batch_t *b)
{
	pthread_mutex_lock(&g_front.mu);
	while (!b->front_done) pthread_cond_wait(&g_front.cv_done, &g_front.mu);
	pthread_mutex_unlock(&g_front.mu);
}

/* fn(ctx, lo, hi) over [0, n) split into nt ranges, on nt - 1 new threads and the caller's */
typedef struct { void (*fn)(void*, int64_t, int64_t); void *ctx; int64_t lo, hi; } par_t;
static void *par_main(void *a)
{
	par_t *p = (par_t*)a;
	p->fn(p->ctx, p->lo, p->hi);
	return 0;
}
static void par_for(int64_t n, int nt, void (*fn)(void*, int64_t, int64_t), void *ctx)
{
	pthread_t th[16];
	par_t a[16];
	int k, started[16];
	if (nt > 16) nt = 16;
	if (nt < 1 || n < 4096) nt = 1;
	for (k = 0; k < nt; ++k) a[k].fn = fn, a[k].ctx = ctx, a[k].lo = n * k / nt, a[k].hi = n * (k + 1) / nt;
	for (k = 1; k < nt; ++k) started[k] = pthread_create(&th[k], 0, par_main, &a[k]) == 0;
	fn(ctx, a[0].lo, a[0].hi);
	for (k = 1; k < nt; ++k) {
		if (started[k]) pthread_join(th[k], 0);
		else fn(ctx, a[k].lo, a[k].hi);                            /* no thread: run the range here */
	}
}

#define READ_THREADS 4

/* the records of a chunk as mb_bseq1_t: names and comments NUL-terminated in the arena, at offset
 * name_off + com_off + 2 i for record i (room for both terminators) */
static void rec_fill(void *ctx, int64_t lo, int64_t hi)
{
	pbatch_t *s = (pbatch_t*)ctx;
	int64_t i;
	for (i = lo; i < hi; ++i) {
		mb_bseq1_t *t = &s->seq[i];
		int64_t nl = s->ch.name_off[i + 1] - s->ch.name_off[i], cl = s->ch.com_off[i + 1] - s->ch.com_off[i];
		t->name = s->arena + s->ch.name_off[i] + s->ch.com_off[i] + 2 * i;
		memcpy(t->name, s->ch.name + s->ch.name_off[i], nl);
		t->name[nl] = 0;
		if (cl > 0) {
			t->comment = t->name + nl + 1;
			memcpy(t->comment, s->ch.comment + s->ch.com_off[i], cl);
			t->comment[cl] = 0;
		}
		t->l_seq = s->ch.seq_off[i + 1] - s->ch.seq_off[i];
		t->seq = (char*)s->ch.seq + s->ch.seq_off[i];
		t->qual = s->ch.has_qual[i] ? (char*)s->ch.qual + s->ch.seq_off[i] : 0;
		t->id = s->id0 + i;
	}
}

/* same[i] (i >= 1): read i shares read i-1's name (mb_qname_same), which joins them in one fragment */
static void rec_same(void *ctx, int64_t lo, int64_t hi)
{
	pbatch_t *s = (pbatch_t*)ctx;
	int64_t i;
	for (i = lo < 1 ? 1 : lo; i < hi; ++i)
		s->same[i] = mb_qname_same(s->seq[i - 1].name, s->seq[i].name);
}

/* the "/1" "/2" trimming of mates' names, and the per-read arrays of the mapping */
static void rec_trim(void *ctx, int64_t lo, int64_t hi)
{
	pbatch_t *s = (pbatch_t*)ctx;
	int64_t i;
	for (i = lo; i < hi; ++i) {
		int32_t j0, j1, l0, l1;
		if (s->seg_cnt[i] != 2) continue;
		j0 = s->seg_off[i], l0 = strlen(s->seq[j0].name);
		j1 = s->seg_off[i] + 1, l1 = strlen(s->seq[j1].name);
		if (l0 < 3 || l0 != l1) continue;
		if (s->seq[j0].name[l0 - 1] != s->seq[j1].name[l1 - 1] && s->seq[j0].name[l0 - 2] == '/')
			s->seq[j0].name[l0 - 2] = s->seq[j1].name[l1 - 2] = 0;
	}
}

/* step 0 (map-main.c:114-173): a mini-batch of reads, then fragments (mates sharing a name) and the
 * trimming of "/1" "/2" name suffixes */
static pbatch_t *read_step(const ref_pipeline_t *p)
{
	const mb_opt_t *opt = p->opt;
	int with_qual = !(opt->flag & MB_F_PAF), with_comment = !!(opt->flag & MB_F_COPY_COMMENT);
	int frag_mode = (p->n_fp > 1 || !!(opt->flag & MB_F_PE));
	pbatch_t *s;
	int32_t i, j;
	int64_t an = 0;
	double t0 = now_s();
	int rc;
	tl_mark(g_pipe.n_batch, "read+");
	if (g_pipe.n_fq == 0) {          /* the first mini-batch opens the native readers on the same files */
		pthread_mutex_lock(&g_mu);
		for (i = 0; i < p->n_fp; ++i) {
			const char *fn = 0;
			for (j = 0; j < g_n_inputs && j < MAX_INPUTS; ++j)
				if (g_inputs[j].fp == p->fp[i]) fn = g_inputs[j].fn;
			if (fn == 0) die("an input was not opened through mb_bseq_open");
			if ((rc = mbg_fq_open(fn, 4, &g_pipe.fq[i])) != MBG_OK) die_rc("mbg_fq_open", rc);
		}
		g_pipe.n_fq = p->n_fp;
		pthread_mutex_unlock(&g_mu);
	}
	s = (pbatch_t*)calloc(1, sizeof(pbatch_t));
	if ((rc = mbg_fq_buf_new(&s->buf)) != MBG_OK) die_rc("mbg_fq_buf_new", rc);
	if (p->n_fp > 1) rc = mbg_fq_read_frag(p->n_fp, g_pipe.fq, p->mb_size, with_qual, with_comment, s->buf, &s->ch);
	else rc = mbg_fq_read(g_pipe.fq[0], p->mb_size, with_qual, with_comment, frag_mode, 40000, opt->max_mb_size, s->buf, &s->ch);
	if (rc != MBG_OK) die_rc("reading the input", rc);
	tl_mark(g_pipe.n_batch, "parsed");
	if (s->ch.status == -2 && p->n_fp == 1) {   /* bseq.c:124-127 */
		if (s->ch.n) {
			int64_t l = s->ch.name_off[s->ch.n] - s->ch.name_off[s->ch.n - 1];
			fprintf(stderr, "[WARNING]\033[1;31m failed to parse the FASTA/FASTQ record next to '%.*s'. Continue anyway.\033[0m\n",
			        (int)l, s->ch.name + s->ch.name_off[s->ch.n - 1]);
		} else fprintf(stderr, "[WARNING]\033[1;31m failed to parse the first FASTA/FASTQ record. Continue anyway.\033[0m\n");
	}
	if (s->ch.n == 0) {
		mbg_fq_buf_free(s->buf);
		free(s);
		return 0;
	}
	s->n_seq = s->ch.n;
	an = s->ch.name_off[s->n_seq] + s->ch.com_off[s->n_seq] + 2 * (int64_t)s->n_seq;
	s->arena = (char*)malloc(an + 1);
	s->seq = (mb_bseq1_t*)calloc(s->n_seq, sizeof(mb_bseq1_t));
	s->n_hit = (int32_t*)calloc(3 * s->n_seq + 3, sizeof(int32_t));
	s->seg_off = s->n_hit + s->n_seq + 1, s->seg_cnt = s->seg_off + s->n_seq + 1;
	s->hit = (mb_hit_t**)calloc(s->n_seq, sizeof(mb_hit_t*));
	if (!s->arena || !s->seq || !s->n_hit || !s->hit) die("out of memory (mini-batch)");
	s->id0 = g_pipe.n_seq;
	par_for(s->n_seq, READ_THREADS, rec_fill, s);
	g_pipe.n_seq += s->n_seq;
	for (i = 0; i < 4; ++i) s->pes[i].failed = 1;
	if (frag_mode) {
		s->same = (uint8_t*)calloc(s->n_seq + 1, 1);
		if (s->same == 0) die("out of memory (mini-batch)");
		par_for(s->n_seq, READ_THREADS, rec_same, s);
	}
	for (i = 1, j = 0; i <= s->n_seq; ++i) {
		if (i == s->n_seq || !frag_mode || !s->same[i]) {
			if (i - j > 2) die("a fragment of more than two reads");      /* the reference asserts the same */
			s->seg_cnt[s->n_frag] = i - j;
			s->seg_off[s->n_frag++] = j;
			if (i - j == 2) s->n_pe++;
			j = i;
		}
	}
	free(s->same);
	s->same = 0;
	if (s->n_pe > 0) par_for(s->n_frag, READ_THREADS, rec_trim, s);
	/* the front pass of this mini-batch, overlapping the previous one's mapping */
	{
		const mb_idx_t *idx = p->idx;
		batch_t *b = &s->b;
		int32_t k;
		s->mt = (uint8_t*)calloc(s->n_seq + 1, 1);
		s->qlen = (int32_t*)malloc((s->n_seq + 1) * sizeof(int32_t));
		s->sq = (const char**)malloc((s->n_seq + 1) * sizeof(char*));
		s->nm = (const char**)malloc((s->n_seq + 1) * sizeof(char*));
		for (k = 0; k < s->n_frag; ++k)       /* worker_for_se_batch's conversion (map-main.c:66) */
			for (j = 0; j < s->seg_cnt[k]; ++j)
				s->mt[s->seg_off[k] + j] = !idx->is_meth ? L2B_METH_NONE : (j & 1) == 0 ? L2B_METH_C2T : L2B_METH_G2A;
		for (k = 0; k < s->n_seq; ++k)
			s->qlen[k] = (int32_t)s->seq[k].l_seq, s->sq[k] = s->seq[k].seq, s->nm[k] = s->seq[k].name, s->tot_len += s->seq[k].l_seq;
		memset(b, 0, sizeof(*b));
		b->opt = opt, b->idx = idx, b->n_seq = s->n_seq, b->mt = s->mt, b->qlen = s->qlen, b->seq = s->sq, b->name = s->nm;
		b->n_hit = s->n_hit, b->hit = s->hit;
		b->seq_no = g_pipe.n_batch;
		if (g_front.started) front_queue(b, (int)(g_pipe.n_batch++ % MBX_FRONT_SLOTS), s->ch.nt4, s->ch.seq_off);
		else {
			front_reads(b, (int)(g_pipe.n_batch++ % MBX_FRONT_SLOTS), s->ch.nt4, s->ch.seq_off);
			b->front_done = 1;
		}
	}
	s->t_read = now_s() - t0;
	tl_mark(s->b.seq_no, "read-");
	return s;
}

/* formatting of fragments [k*FMT_FRAGS, (k+1)*FMT_FRAGS) as the output step does (map-main.c:207-246) */
#define FMT_FRAGS 2048
typedef struct { pbatch_t *s; const mb_opt_t *opt; const mb_idx_t *idx; } fmt_ctx_t;

/* One fragment (n_seg segments t[0..n_seg)) written as the output step writes it (map-main.c:213-240):
 * each hit kept by the secondary filter (out_n, out_s), or the unmapped record; then its hits freed. */
void mbg_map_fmt_fragment(kstring_t *out, const l2b_t *l2b, const mb_opt_t *opt, mb_bseq1_t *t, int32_t n_seg, int32_t *n_hit,
                          mb_hit_t **hit)
{
	int32_t i, j;
	void *km = 0;                /* libc allocation (the CLI's default build, MB_F_NO_KALLOC) */
	for (i = 0; i < n_seg; ++i) {
		int32_t mate_qlen = 0;
		if (n_seg > 1) mate_qlen = t[i != n_seg - 1 ? i + 1 : 0].l_seq;
		if (n_hit[i] > 0) {
			int32_t n_sec = 0;
			for (j = 0; j < n_hit[i]; ++j) {
				const mb_hit_t *h = &hit[i][j];
				if (h->parent == h->id || n_sec < opt->out_n) {
					if (h->parent != h->id) {
						const mb_hit_t *q = &hit[i][h->parent];
						if (q->p && h->p) {
							if (h->p->dp_max < (double)opt->out_s * q->p->dp_max) continue;
						} else {
							if (h->score < (double)opt->out_s * q->score) continue;
						}
					}
					fmt_one(km, out, l2b, &t[i], n_seg, n_hit, hit, j, opt, i, mate_qlen);
					n_sec += (h->parent != h->id);
				}
			}
		} else if (!(opt->flag & MB_F_NO_UNMAP)) {
			fmt_one(km, out, l2b, &t[i], n_seg, n_hit, hit, -1, opt, i, mate_qlen);
		}
	}
	for (i = 0; i < n_seg; ++i) {
		for (j = 0; j < n_hit[i]; ++j) free(hit[i][j].p);
		free(hit[i]);
		hit[i] = 0;
	}
}

static void fmt_task(void *data, int64_t k, int tid)
{
	const fmt_ctx_t *c = (const fmt_ctx_t*)data;
	pbatch_t *s = c->s;
	int32_t f, f1 = (int32_t)((k + 1) * FMT_FRAGS < s->n_frag ? (k + 1) * FMT_FRAGS : s->n_frag);
	(void)tid;
	for (f = (int32_t)(k * FMT_FRAGS); f < f1; ++f) {
		int32_t seg_st = s->seg_off[f];
		mbg_map_fmt_fragment(&s->out[k], c->idx->l2b, c->opt, &s->seq[seg_st], s->seg_cnt[f], &s->n_hit[seg_st], &s->hit[seg_st]);
	}
}

/* step 1 (map-main.c:176-192): single-end mapping, the insert-size model, pairing */
static void *map_step(pbatch_t *s)
{
	const ref_pipeline_t *p = g_pipe.p;
	const mb_opt_t *opt = p->opt;
	batch_t *b = &s->b;
	int32_t k, *pair_off = 0, n_pair = 0;
	double t0 = now_s();
	tl_mark(b->seq_no, "map+");
	front_wait(b);
	se_reads(b, opt->n_thread);
	if ((opt->flag & MB_F_PE) && s->n_frag < s->n_seq && !(opt->flag & (MB_F_NO_PAIRING | MB_F_NO_ALN))) { /* map-main.c:180-192 */
		if ((opt->flag & MB_F_PE_PREDEF) || s->n_pe < 20) {
			s->pes[1].failed = 0;
			s->pes[1].avg = opt->pe_avg, s->pes[1].std = opt->pe_std;
			s->pes[1].lo = opt->pe_lo, s->pes[1].hi = opt->pe_hi;
		} else {
			tl_mark(b->seq_no, "pestat+");
			mbg_pestat(opt, s->n_frag, s->seg_off, s->seg_cnt, s->n_hit, s->hit, s->pes, opt->n_thread);
		}
		pair_off = (int32_t*)malloc((s->n_frag + 1) * sizeof(int32_t));
		for (k = 0; k < s->n_frag; ++k)
			if (s->seg_cnt[k] == 2) pair_off[n_pair++] = s->seg_off[k];
		b->n_pair = n_pair, b->pair_off = pair_off, b->pes = s->pes;
		pair_reads(b);
	}
	map_free(b);
	free(pair_off);
	s->t_map = now_s() - t0;
	tl_mark(b->seq_no, "map-");
	return s;
}

/* step 2: the formatted records in order, then the mini-batch's memory (map-main.c:242-256) */
static void write_step(pbatch_t *s)
{
	const ref_pipeline_t *p = g_pipe.p;
	int32_t k;
	fmt_ctx_t fc;
	double t0 = now_s(), t1;
	s->n_out = (s->n_frag + FMT_FRAGS - 1) / FMT_FRAGS;
	s->out = (kstring_t*)calloc(s->n_out + 1, sizeof(kstring_t));
	fc.s = s, fc.opt = p->opt, fc.idx = p->idx;
	tl_mark(s->b.seq_no, "fmt+");
	mbs_run(g_fmt_pool, s->n_out, 1, fmt_task, 0, &fc, 0);
	t1 = now_s();
	s->t_fmt = t1 - t0;
	t0 = t1;
	for (k = 0; k < s->n_out; ++k) {
		if (s->out[k].l) fwrite(s->out[k].s, 1, s->out[k].l, p->fp_out);
		free(s->out[k].s);
	}
	tl_mark(s->b.seq_no, "write-");
	if (kom_verbose >= 3)
		fprintf(stderr, "[M::%s::%.3f*%.2f] mapped %ld bp in %ld sequences\n", "worker_pipeline", kom_realtime(), kom_percent_cpu(),
		        (long)s->tot_len, (long)s->n_seq);
	if (g_stats)
		fprintf(stderr, "[mbg_map] mini-batch of %d reads: read %.3f s, map %.3f s, format %.3f s, write %.3f s\n", s->n_seq, s->t_read,
		        s->t_map, s->t_fmt, now_s() - t0);
	free(s->out); free(s->hit); free(s->n_hit); free(s->seq); free(s->arena);
	free(s->mt); free(s->qlen); free(s->sq); free(s->nm);
	mbg_fq_buf_free(s->buf);
	free(s);
}

static void *pipe_step(void *data, int step, void *in)
{
	(void)data;
	if (step == 0) return read_step(g_pipe.p);
	if (step == 1) return map_step((pbatch_t*)in);
	write_step((pbatch_t*)in);
	return 0;
}

void __wrap_kt_pipeline(int n_threads, void *(*func)(void*, int, void*), void *shared_data, int n_steps)
{
	const ref_pipeline_t *p = (const ref_pipeline_t*)shared_data;
	int i;
	if (n_steps != 3) {          /* not the mapper's three-step pipeline */
		__real_kt_pipeline(n_threads, func, shared_data, n_steps);
		return;
	}
	if (p->opt == 0 || p->idx == 0 || p->mb_size != p->opt->mb_size || p->n_fp < 1 || p->n_fp > MAX_INPUTS)
		die("the pipeline record does not match minibwa 0.7-r421's layout");
	pool_for(p->opt->n_thread);
	if (g_fmt_pool == 0) {       /* formatting overlaps the next mini-batch's mapping */
		const char *e = getenv("MBG_FMT_THREADS");
		int n = e ? atoi(e) : p->opt->n_thread;
		g_fmt_pool = mbs_pool_init(n > 0 ? n : 1, 64 * 1024);
	}
	mbg_map_fmt_init(p->opt);
	g_pipe.p = p, g_pipe.n_fq = 0, g_pipe.n_seq = p->n_seq, g_pipe.n_batch = 0;
	{
		const char *e = getenv("MBG_TIMELINE");
		g_tl = e && atoi(e) > 0;
		g_tl_t0 = now_s();
		atomic_store(&g_tl_n, 0);
	}
	{
		const char *e = getenv("MBG_PIPE_DEPTH");
		g_pipe_depth = e ? atoi(e) : 4;
		if (g_pipe_depth < 3 || n_threads < 3) g_pipe_depth = 3;
		if (g_pipe_depth > MBX_FRONT_SLOTS) g_pipe_depth = MBX_FRONT_SLOTS;
	}
	if (!g_front.started) {
		const char *e = getenv("MBG_FRONT_THREAD"), *e2 = getenv("MBG_FRONT_THREADS");
		int n = e2 ? atoi(e2) : g_pipe_depth - 1;
		if (e && atoi(e) == 0) n = 0;
		if (n > FRONT_MAX_THREADS) n = FRONT_MAX_THREADS;
		if (n > g_pipe_depth - 1) n = g_pipe_depth - 1;    /* fronts waiting at once: the mini-batches in flight but one */
		if (n > 0) {
			int k;
			g_front.stop = 0, g_front.n_th = n;
			for (k = 0; k < n; ++k)
				if (pthread_create(&g_front.th[k], 0, front_main, 0) != 0) die("pthread_create failed (front thread)");
			g_front.started = 1;
		}
	}
	__real_kt_pipeline(n_threads >= 3 ? g_pipe_depth : n_threads, pipe_step, 0, n_steps);
	if (g_front.started) {        /* every queued front has been waited for by its map step */
		pthread_mutex_lock(&g_front.mu);
		g_front.stop = 1;
		pthread_cond_broadcast(&g_front.cv_q);
		pthread_mutex_unlock(&g_front.mu);
		for (i = 0; i < g_front.n_th; ++i) pthread_join(g_front.th[i], 0);
		g_front.started = 0;
	}
	for (i = 0; i < g_pipe.n_fq; ++i) mbg_fq_close(g_pipe.fq[i]);
	g_pipe.n_fq = 0;
	if (g_tl) {
		int n = atomic_load(&g_tl_n);
		if (n > TL_CAP) n = TL_CAP;
		for (i = 0; i < n; ++i)
			fprintf(stderr, "[mbg_map] timeline %8.1f ms  batch %ld %s\n", 1e3 * (g_tl_ev[i].t - g_tl_t0), (long)g_tl_ev[i].batch, g_tl_ev[i].ev);
	}
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

/* ---- groups: reads mapped by tasks of an open run (the serving loop) ------------------------------ */

struct mbg_group_s {
	batch_t b;
	uint8_t *mt;
	mb_pestat_t pes[4];
	int pair;                     /* tasks pair their two reads (mb_pair) */
};

mbg_group_t *mbg_group_new(const mb_opt_t *opt, const mb_idx_t *idx, int32_t n_seq, const int32_t *qlen, const char *const *seq,
                           const char *const *name, int slot)
{
	int32_t i, is_pe = !!(opt->flag & MB_F_PE), is_meth = !!(opt->flag & MB_F_METH);
	mbg_group_t *g = (mbg_group_t*)calloc(1, sizeof(mbg_group_t));
	if (g == 0 || n_seq <= 0 || (is_meth && !idx->is_meth)) {
		free(g);
		return 0;
	}
	g->mt = (uint8_t*)calloc(n_seq, 1);
	g->b.n_hit = (int32_t*)calloc(n_seq, sizeof(int32_t));
	g->b.hit = (mb_hit_t**)calloc(n_seq, sizeof(mb_hit_t*));
	if (!g->mt || !g->b.n_hit || !g->b.hit) die("out of memory (group)");
	for (i = 0; i < n_seq; ++i)        /* mb_map_batch's conversion (map-algo.c:733-737) */
		g->mt[i] = !is_meth ? L2B_METH_NONE : !is_pe || (i & 1) == 0 ? L2B_METH_C2T : L2B_METH_G2A;
	g->b.opt = opt, g->b.idx = idx, g->b.n_seq = n_seq, g->b.qlen = qlen, g->b.seq = (const char *const*)seq;
	g->b.name = (const char *const*)name, g->b.mt = g->mt;
	for (i = 0; i < 4; ++i) g->pes[i].failed = 1;       /* mb_map_batch's predefined statistics (map-algo.c:755-766) */
	g->pes[1].failed = 0;
	g->pes[1].avg = opt->pe_avg, g->pes[1].std = opt->pe_std;
	g->pes[1].lo = opt->pe_lo, g->pes[1].hi = opt->pe_hi;
	g->b.pes = g->pes;
	g->pair = is_pe && !(opt->flag & (MB_F_NO_PAIRING | MB_F_NO_ALN));
	front_reads(&g->b, slot, 0, 0);
	g->b.front_done = 1;
	return g;
}

void mbg_group_map(mbg_group_t *g, int32_t r0, int32_t n)
{
	const int tid = mbs_tid();
	if (tid < 0) die("mbg_group_map outside a task");
	se_task(&g->b, r0, tid);
	if (n == 2) {
		se_task(&g->b, r0 + 1, tid);
		if (g->pair) {
			int32_t len[2];
			char *sq[2];
			len[0] = g->b.qlen[r0], len[1] = g->b.qlen[r0 + 1];
			sq[0] = (char*)g->b.seq[r0], sq[1] = (char*)g->b.seq[r0 + 1];
			mb_pair(mb_tbuf_km(g_tbuf[tid]), g->b.opt, g->b.idx->l2b, &g->b.n_hit[r0], &g->b.hit[r0], g->pes, len, sq);
		}
	}
}

int32_t *mbg_group_n_hit(mbg_group_t *g) { return g->b.n_hit; }
mb_hit_t **mbg_group_hit(mbg_group_t *g) { return g->b.hit; }

void mbg_group_free(mbg_group_t *g)
{
	if (g == 0) return;
	map_free(&g->b);
	free(g->b.n_hit); free(g->b.hit); free(g->mt);
	free(g);
}

mbs_pool_t *mbg_map_pool(int n_threads) { return pool_for(n_threads); }
void mbg_map_round(void *data, void **reqs, int64_t n) { round_cb(data, reqs, n); }
int64_t mbg_map_flush(void) { return g_flush; }

/* ---- index --------------------------------------------------------------------------------------- */

/* The autotuner's choices (bench.py --autotune): KEY=VALUE lines of MBG_ settings in the file named by
 * MBG_TUNE, or else in tuned.env beside the program's directory (minibwa_amd/tuned.env for the
 * programs in minibwa_amd/bin); a setting already in the environment wins.  MBG_TUNE=none reads none.
 * The settings change launch shapes and batch sizes, never results. */
static void tune_load(void)
{
	static int done;
	char path[4096], line[512];
	const char *e = getenv("MBG_TUNE");
	FILE *fp;
	if (done) return;
	done = 1;
	if (e && strcmp(e, "none") == 0) return;
	if (e && *e) snprintf(path, sizeof(path), "%s", e);
	else {
		ssize_t n = readlink("/proc/self/exe", path, sizeof(path) - 32);
		char *sl;
		if (n <= 0) return;
		path[n] = 0;
		if ((sl = strrchr(path, '/')) == 0) return;
		strcpy(sl, "/../tuned.env");
	}
	if ((fp = fopen(path, "r")) == 0) return;
	while (fgets(line, sizeof(line), fp)) {
		char *eq = strchr(line, '='), *v, *nl;
		if (line[0] == '#' || eq == 0 || strncmp(line, "MBG_", 4) != 0) continue;
		*eq = 0, v = eq + 1;
		if ((nl = strpbrk(v, "\r\n")) != 0) *nl = 0;
		setenv(line, v, 0);
	}
	fclose(fp);
}

mb_idx_t *__real_mb_idx_load(const char *prefix, int32_t is_meth);
mb_idx_t *__wrap_mb_idx_load(const char *prefix, int32_t is_meth)
{
	mb_idx_t *idx;
	pthread_mutex_lock(&g_mu);
	tune_load();
	pthread_mutex_unlock(&g_mu);
	idx = __real_mb_idx_load(prefix, is_meth);
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
/* The reference bases around the first anchor of each of a read's first chains, towards the caches:
 * the read's extension reads them (l2b_getseq) a few microseconds later, and the first touch of a
 * random spot of a multi-gigabase reference is otherwise a full memory miss on the mapping thread.
 * u[i]'s low 32 bits count chain i's anchors, which follow each other in a. */
static void prefetch_targets(const l2b_t *l2b, int32_t n_u, const uint64_t *u, const mb_anchor_t *a)
{
	int32_t i;
	int64_t k = 0;
	for (i = 0; i < n_u && i < 4; ++i) {
		const mb_anchor_t *x = &a[k];
		const int64_t tid = x->sid >> 1;
		if (tid >= 0 && (uint64_t)tid < l2b->n_ctg) {
			const int64_t g = l2b->ctg[tid].off + x->tpos;
			const char *p = (const char*)l2b->pac + (g >> 2);
			const char *lo = (const char*)l2b->pac + (l2b->ctg[tid].off >> 2), *hi = lo + ((l2b->ctg[tid].len + 3) >> 2);
			int d;
			for (d = -192; d <= 192; d += 64)
				if (p + d >= lo && p + d < hi) __builtin_prefetch(p + d);
		}
		k += (uint32_t)u[i];
	}
}

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
			prefetch_targets(l2b, n_u, u, c);
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
	/* the CIGAR lands in ez's own buffer, grown first to the most a job can write with the allocator
	 * ksw2 grows it with (kalloc's; libc's for a null km); what it held is overwritten, so not copied */
	if (ez->m_cigar < r.qlen + r.tlen + 1) {
		int m = ez->m_cigar * 2 > r.qlen + r.tlen + 1 ? ez->m_cigar * 2 : r.qlen + r.tlen + 1;
		kfree(km, ez->cigar);
		ez->cigar = (uint32_t*)kmalloc(km, (size_t)m * 4);
		if (ez->cigar == 0) die("out of memory (CIGAR)");
		ez->m_cigar = m;
	}
	r.cigar = ez->cigar;
	if (solo) {
		mbx_req_t *p = &r;
		pthread_mutex_lock(&g_solo_mu);
		mbx_run_dp(dual, 1, &p);
		pthread_mutex_unlock(&g_solo_mu);
	} else post_and_wait(&r);
	ez->max = (uint32_t)r.dp.max & 0x7fffffff, ez->zdropped = r.dp.zdropped ? 1 : 0;
	ez->max_q = r.dp.max_q, ez->max_t = r.dp.max_t;
	ez->mqe = r.dp.mqe, ez->mqe_t = r.dp.mqe_t;
	ez->mte = r.dp.mte, ez->mte_q = r.dp.mte_q;
	ez->score = r.dp.score, ez->reach_end = r.dp.reach_end;
	ez->n_cigar = r.dp.n_cigar;
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
