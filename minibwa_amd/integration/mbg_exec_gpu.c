/*
 * mbg_exec_gpu.c -- the batched mapper's device requests (mbg_exec.h) executed by libmbg.so on one
 * MI355X: the front pass through mbg_map_front (on one of MBX_FRONT_SLOTS contexts that share the
 * resident index, so that a mini-batch's front overlaps the previous one's rounds), each round's DP
 * jobs through mbg_extd2_batch /
 * mbg_extz2_batch (one call per scoring matrix), its local-SW jobs through mbg_ll_batch.  Sequences
 * are packed into page-locked staging buffers so each round crosses PCIe in a few large copies.
 * Any library error aborts with its message, as the reference does on internal failures.
 *
 * Environment: MBG_DEVICE (HIP device, default 0), MBG_PRE_LEN (k-mer interval cache length on the
 * device; default 15 for indexes over 2^30 rows, else 12; results do not depend on it),
 * MBG_COUNT_WORK=N (count SMEM pass 1's algorithmic work during the first N front passes; counting
 * adds atomics to the seeding kernels), MBG_MAP_STATS=1 (one stderr line per front pass with its
 * device times).
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "kalloc.h"
#include "mbg_exec.h"

typedef char sai_layout_check[sizeof(mb_sai_t) == sizeof(mbg_sai_t) ? 1 : -1];
typedef char anchor_layout_check[sizeof(mb_anchor_t) == sizeof(mbg_anchor_t) ? 1 : -1];

static mbg_ctx_t *g_ctx;                        /* the index, the DP and local-SW rounds */
static mbg_ctx_t *g_fctx[MBX_FRONT_SLOTS];      /* front passes: one context per slot, on g_ctx's index */
static const mb_idx_t *g_idx;
static mbx_stats_t g_st;
static pthread_mutex_t g_st_mu = PTHREAD_MUTEX_INITIALIZER;
static int g_front_log;          /* MBG_MAP_STATS: one line per front pass */
static double g_pack_s;          /* host time packing DP rounds */
static long g_count_fronts;      /* MBG_COUNT_WORK=N: count pass-1 work during the first N front passes */

/* staging buffers (page-locked, grow-only) */
typedef struct { void *p; size_t n; } hbuf_t;
static hbuf_t b_q, b_t, b_jobs, b_res, b_cig, b_ll_jobs, b_ll_res;

static void chk(int rc, const char *what)
{
	if (rc != MBG_OK) {
		fprintf(stderr, "[mbg_exec] %s failed: %s (%s)\n", what, mbg_strerror(rc), mbg_last_hip_error());
		abort();
	}
}

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void *grow(hbuf_t *b, size_t bytes)
{
	if (bytes == 0) bytes = 16;
	if (b->n < bytes) {
		size_t n = bytes + bytes / 2 + 4096;
		if (b->p) chk(mbg_host_free(g_ctx, b->p), "mbg_host_free");
		chk(mbg_host_alloc(g_ctx, n, &b->p), "mbg_host_alloc");
		b->n = n;
	}
	return b->p;
}

const char *mbx_name(void) { return "gpu"; }

void mbx_index(const mb_idx_t *idx)
{
	const mb_bwt_t *bwt = idx->bwt;
	const l2b_t *l = idx->l2b;
	mbg_bwt_desc_t d;
	mbg_l2b_desc_t ld;
	uint64_t *len, i;
	const char *s;
	int pre;
	if (g_ctx == 0) {
		s = getenv("MBG_DEVICE");
		chk(mbg_init(s ? atoi(s) : 0, &g_ctx), "mbg_init");
	}
	s = getenv("MBG_COUNT_WORK");
	if (s && atoi(s) > 0) g_count_fronts = atol(s);
	s = getenv("MBG_MAP_STATS");
	g_front_log = s && atoi(s);
	s = getenv("MBG_PRE_LEN");
	pre = s ? atoi(s) : bwt->seq_len > (1ULL << 30) ? 15 : 12;
	d.primary = bwt->primary;
	memcpy(d.L2, bwt->L2, sizeof(d.L2));
	d.seq_len = bwt->seq_len, d.data = bwt->data, d.sa_bit = bwt->sa_bit, d.n_sa = bwt->n_sa, d.sa = bwt->sa;
	chk(mbg_bwt_upload(g_ctx, &d, pre), "mbg_bwt_upload");
	len = (uint64_t*)malloc(sizeof(uint64_t) * (l->n_ctg ? l->n_ctg : 1));
	for (i = 0; i < l->n_ctg; ++i) len[i] = l->ctg[i].len;
	ld.tot_len = l->tot_len, ld.n_ctg = l->n_ctg, ld.ctg_len = len;
	ld.n_pac = l->n_pac, ld.pac = l->pac, ld.n_ambi = l->n_ambi, ld.ambi = (const uint64_t*)l->ambi;
	chk(mbg_l2b_upload(g_ctx, &ld), "mbg_l2b_upload");
	free(len);
	g_idx = idx;
}

static mbg_ctx_t *front_ctx(int slot)
{
	mbg_ctx_t *c;
	if (slot < 0 || slot >= MBX_FRONT_SLOTS) slot = 0;
	pthread_mutex_lock(&g_st_mu);
	if (g_fctx[slot] == 0) {
		const char *s = getenv("MBG_DEVICE");
		chk(mbg_init(s ? atoi(s) : 0, &g_fctx[slot]), "mbg_init");
		chk(mbg_idx_attach(g_fctx[slot], g_ctx), "mbg_idx_attach");
		if (g_count_fronts > 0) chk(mbg_count_work(g_fctx[slot], 1), "mbg_count_work");
	}
	c = g_fctx[slot];
	pthread_mutex_unlock(&g_st_mu);
	return c;
}

void mbx_front(const mbx_front_in_t *in, mbx_front_t *out)
{
	mbg_front_opt_t o;
	mbg_front_t r;
	mbg_ctx_t *c = front_ctx(in->slot);
	double t0 = now_s();
	int k;
	o.min_len = in->min_len, o.max_sub_occ = in->max_sub_occ, o.max_occ = in->max_occ;
	o.n_copt = in->n_copt, o.copt = in->copt, o.copt_idx = in->copt_idx, o.no_chain = in->no_chain;
	chk(mbg_map_front(c, in->n, in->nt4, in->off, in->len, &o, &r), "mbg_map_front");
	out->n = in->n;
	out->intv_off = r.intv_off, out->n_intv = r.n_intv, out->intv = (const mb_sai_t*)r.intv;
	out->anc_off = r.anc_off, out->n_anc = r.n_anc, out->anc = (const mb_anchor_t*)r.anc;
	out->n_u = r.n_u, out->u = r.u, out->n_chn = r.n_chn, out->chn = (const mb_anchor_t*)r.chn;
	out->dev_ms = r.ms_total;
	pthread_mutex_lock(&g_st_mu);
	++g_st.fronts, g_st.front_reads += in->n, g_st.front_s += now_s() - t0;
	if (g_front_log)
		fprintf(stderr, "[mbg_exec] front %ld reads %d records %ld slots %ld seed1_ms %.4f seed2_ms %.4f group_ms %.4f anchor_ms %.4f "
		        "chain_ms %.4f copy_ms %.4f total_ms %.4f seed1_occ_blocks %lu seed1_kmer_lookups %lu\n", (long)g_st.fronts, in->n,
		        (long)r.n_rec, (long)r.n_slots, r.ms_seed1, r.ms_seed2, r.ms_group, r.ms_anchor, r.ms_chain, r.ms_copy, r.ms_total,
		        (unsigned long)r.seed1_occ_blocks, (unsigned long)r.seed1_kmer_lookups);
	if (g_count_fronts > 0 && --g_count_fronts == 0)
		for (k = 0; k < MBX_FRONT_SLOTS; ++k)
			if (g_fctx[k]) chk(mbg_count_work(g_fctx[k], 0), "mbg_count_work");
	pthread_mutex_unlock(&g_st_mu);
}

/* the requests of one scoring matrix in one library call */
static void dp_group(int dual, int32_t n, mbx_req_t *const *req, const int32_t *sel)
{
	int64_t qb = 0, tb = 0, cb = 0;
	int32_t i;
	uint8_t *q, *t;
	mbg_dp_job_t *jobs;
	mbg_dp_result_t *res;
	uint32_t *cig;
	for (i = 0; i < n; ++i) {
		const mbx_req_t *r = req[sel[i]];
		qb += r->qlen, tb += r->tlen, cb += (int64_t)r->qlen + r->tlen;
	}
	double t0 = now_s();
	q = (uint8_t*)grow(&b_q, qb + 16), t = (uint8_t*)grow(&b_t, tb + 16);
	jobs = (mbg_dp_job_t*)grow(&b_jobs, (size_t)n * sizeof(mbg_dp_job_t));
	res = (mbg_dp_result_t*)grow(&b_res, (size_t)n * sizeof(mbg_dp_result_t));
	cig = (uint32_t*)grow(&b_cig, (size_t)cb * 4 + 16);
	/* one thread packs: the mapper's workers keep every core busy meanwhile, and a parallel region
	 * would wait for all of its threads to be scheduled */
	for (i = 0, qb = tb = cb = 0; i < n; ++i) {
		const mbx_req_t *r = req[sel[i]];
		mbg_dp_job_t *j = &jobs[i];
		j->qoff = qb, j->toff = tb, j->qlen = r->qlen, j->tlen = r->tlen;
		j->q = r->gq, j->e = r->ge, j->q2 = r->gq2, j->e2 = r->ge2;
		j->w = r->w, j->zdrop = r->zdrop, j->end_bonus = r->end_bonus, j->flag = r->flag;
		j->cigar_off = cb;
		if (r->qlen) memcpy(q + qb, r->q, r->qlen);
		if (r->tlen) memcpy(t + tb, r->t, r->tlen);
		qb += r->qlen, tb += r->tlen, cb += (int64_t)r->qlen + r->tlen;
	}
	g_pack_s += now_s() - t0;
	if (dual) chk(mbg_extd2_batch(g_ctx, n, jobs, q, qb, t, tb, req[sel[0]]->m, req[sel[0]]->mat, res, cig, cb), "mbg_extd2_batch");
	else chk(mbg_extz2_batch(g_ctx, n, jobs, q, qb, t, tb, req[sel[0]]->m, req[sel[0]]->mat, res, cig, cb), "mbg_extz2_batch");
	for (i = 0; i < n; ++i) {               /* each CIGAR into its poster's buffer (room for qlen + tlen) */
		mbx_req_t *r = req[sel[i]];
		r->dp = res[i];
		if (r->dp.cells < 0 || r->dp.n_cigar < 0 || r->dp.n_cigar > r->qlen + r->tlen + 1) {
			fprintf(stderr, "[mbg_exec] a DP job returned %d CIGAR operations for %d + %d bases\n", r->dp.n_cigar, r->qlen, r->tlen);
			abort();
		}
		if (r->dp.n_cigar) memcpy(r->cigar, cig + jobs[i].cigar_off, (size_t)r->dp.n_cigar * 4);
	}
}

/* Requests are grouped by their scoring matrix (and, for local SW, by the shared gap and xtra
 * arguments): a library call takes one of each. */
static int32_t *g_sel, *g_grp;
static int64_t g_sel_m;

static void sel_buf(int64_t n)
{
	if (n > g_sel_m) {
		g_sel_m = n + 1024;
		g_sel = (int32_t*)realloc(g_sel, g_sel_m * sizeof(int32_t));
		g_grp = (int32_t*)realloc(g_grp, g_sel_m * sizeof(int32_t));
		if (g_sel == 0 || g_grp == 0) { fprintf(stderr, "[mbg_exec] out of memory\n"); abort(); }
	}
}

void mbx_run_dp(int dual, int32_t n, mbx_req_t *const *req)
{
	int32_t i, g, n_grp = 0, n_sel;
	double t0 = now_s();
	sel_buf(n);
	for (i = 0; i < n; ++i) g_grp[i] = -1;
	for (i = 0; i < n; ++i) {          /* group ids in order of first appearance */
		int32_t j;
		if (g_grp[i] >= 0) continue;
		g_grp[i] = n_grp;
		for (j = i + 1; j < n; ++j)
			if (g_grp[j] < 0 && req[j]->m == req[i]->m && memcmp(req[j]->mat, req[i]->mat, (size_t)req[i]->m * req[i]->m) == 0)
				g_grp[j] = n_grp;
		++n_grp;
	}
	for (g = 0; g < n_grp; ++g) {
		if (n_grp == 1) {
			for (i = 0; i < n; ++i) g_sel[i] = i;
			n_sel = n;
		} else
			for (i = 0, n_sel = 0; i < n; ++i)
				if (g_grp[i] == g) g_sel[n_sel++] = i;
		dp_group(dual, n_sel, req, g_sel);
	}
	pthread_mutex_lock(&g_st_mu);
	++g_st.dp_calls, g_st.dp_jobs += n, g_st.dp_s += now_s() - t0;
	pthread_mutex_unlock(&g_st_mu);
}

void mbx_run_ll(int32_t n, mbx_req_t *const *req)
{
	int32_t i, n_sel, done = 0, first = 0, *sel;
	uint8_t *taken = (uint8_t*)calloc(n, 1);
	double t0 = now_s();
	sel_buf(n);
	sel = g_sel;
	while (done < n) {
		const mbx_req_t *k;
		int64_t qb = 0, tb = 0;
		uint8_t *q, *t;
		mbg_ll_job_t *jobs;
		mbg_ll_result_t *res;
		while (taken[first]) ++first;
		k = req[first];
		for (i = first, n_sel = 0; i < n; ++i) {
			const mbx_req_t *r = req[i];
			if (!taken[i] && r->m == k->m && r->gapo == k->gapo && r->gape == k->gape && r->xtra == k->xtra &&
			    memcmp(r->mat, k->mat, (size_t)k->m * k->m) == 0)
				sel[n_sel++] = i, taken[i] = 1, qb += r->qlen, tb += r->tlen;
		}
		done += n_sel;
		q = (uint8_t*)grow(&b_q, qb + 16), t = (uint8_t*)grow(&b_t, tb + 16);
		jobs = (mbg_ll_job_t*)grow(&b_ll_jobs, (size_t)n_sel * sizeof(mbg_ll_job_t));
		res = (mbg_ll_result_t*)grow(&b_ll_res, (size_t)n_sel * sizeof(mbg_ll_result_t));
		for (i = 0, qb = tb = 0; i < n_sel; ++i) {
			const mbx_req_t *r = req[sel[i]];
			mbg_ll_job_t *j = &jobs[i];
			memset(j, 0, sizeof(*j));
			j->qoff = qb, j->toff = tb, j->qlen = r->qlen, j->tlen = r->tlen, j->size = r->size;
			if (r->qlen) memcpy(q + qb, r->q, r->qlen);
			if (r->tlen) memcpy(t + tb, r->t, r->tlen);
			qb += r->qlen, tb += r->tlen;
		}
		chk(mbg_ll_batch(g_ctx, n_sel, jobs, q, qb, t, tb, k->m, k->mat, k->gapo, k->gape, k->xtra, res), "mbg_ll_batch");
		for (i = 0; i < n_sel; ++i) req[sel[i]]->ll = res[i];
	}
	free(taken);
	pthread_mutex_lock(&g_st_mu);
	++g_st.ll_calls, g_st.ll_jobs += n, g_st.ll_s += now_s() - t0;
	pthread_mutex_unlock(&g_st_mu);
}

void mbx_sa(int64_t n, uint64_t *x)
{
	chk(mbg_sa_batch(g_ctx, n, x), "mbg_sa_batch");
}

/* intervals appended into the caller's containers with kalloc on its km, as mb_seed_intv_batch does
 * (bwt.c:368-371) */
void mbx_seed(void *km, int32_t n, const int32_t *len, uint8_t *const *seq, int32_t min_len, int32_t max_sub_occ, mb_sai_v *v)
{
	mbg_sai_t *out = 0;
	int64_t *off = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
	int32_t i;
	chk(mbg_seed_batch(g_ctx, n, len, (const uint8_t *const*)seq, min_len, max_sub_occ, &out, off), "mbg_seed_batch");
	for (i = 0; i < n; ++i) {
		int64_t k, m = off[i + 1] - off[i];
		v[i].n = 0;
		if ((size_t)m > v[i].m) {
			v[i].m = (size_t)m + ((size_t)m >> 1) + 16;
			v[i].a = (mb_sai_t*)krealloc(km, v[i].a, v[i].m * sizeof(mb_sai_t));   /* kalloc's, as Kgrow */
		}
		for (k = 0; k < m; ++k) memcpy(&v[i].a[v[i].n++], &out[off[i] + k], sizeof(mb_sai_t));
	}
	free(out);
	free(off);
}

void mbx_get_stats(mbx_stats_t *st)
{
	int k;
	pthread_mutex_lock(&g_st_mu);
	*st = g_st;
	st->dp_pack_s = g_pack_s;
	memset(&st->dev, 0, sizeof(st->dev));
	if (g_ctx) chk(mbg_counters(g_ctx, &st->dev, 0), "mbg_counters");
	for (k = 0; k < MBX_FRONT_SLOTS; ++k) {          /* the front passes' counters live on their contexts */
		mbg_counters_t f;
		if (g_fctx[k] == 0) continue;
		chk(mbg_counters(g_fctx[k], &f, 0), "mbg_counters");
		st->dev.fronts += f.fronts, st->dev.front_reads += f.front_reads, st->dev.front_records += f.front_records;
		st->dev.front_slots += f.front_slots, st->dev.ms_seed1 += f.ms_seed1, st->dev.ms_seed2 += f.ms_seed2;
		st->dev.ms_group += f.ms_group, st->dev.ms_anchor += f.ms_anchor, st->dev.ms_chain += f.ms_chain;
		st->dev.ms_front_copy += f.ms_front_copy, st->dev.ms_front += f.ms_front;
		st->dev.seed1_occ_blocks += f.seed1_occ_blocks, st->dev.seed1_kmer_lookups += f.seed1_kmer_lookups;
		st->dev.graph_launches += f.graph_launches, st->dev.graph_instantiations += f.graph_instantiations;
		st->dev.direct_chains += f.direct_chains;
	}
	pthread_mutex_unlock(&g_st_mu);
}

void mbx_mem(int64_t *work, int64_t *dev_free, int64_t *dev_total)
{
	*work = *dev_free = *dev_total = 0;
	chk(mbg_work_bytes(work, dev_free, dev_total), "mbg_work_bytes");
}
