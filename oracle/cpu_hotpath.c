/*
 * cpu_hotpath.c -- times the reference's own hot-path functions on the host cores (the CPU
 * baseline of bench.py).  TEST/BASELINE INFRASTRUCTURE: built into oracle/_ref/ against the
 * compiled reference, never part of libmbg.
 *
 *   cpu_hotpath <idx_prefix> <reads.bin> <jobs.bin> <n_threads>
 *
 * An idx_prefix of "-" skips the index and phase 1 (DP-only timing; reads.bin may hold 0 reads).
 * Phase 1, per sub-batch of 48 reads (sb_seq=24 fragments, map-main.c:163-174) on a work-stealing
 * counter like kt_for: mb_seed_intv_batch (seed.c:56-103), then mb_anchor per read (seed.c:268-339:
 * interval sort/dedup, SA locate through mb_bwt_sa_batch, contig mapping, anchor sort/dedup).
 * Phase 2: ksw_extd2_sse (ksw2_extd2_sse.c:27) -- or ksw_extz2_sse when q==q2 and e==e2 -- on
 * every extension job in jobs.bin (the jobs the GPU pass emitted for the same reads).
 * Prints one JSON line with the wall time of each phase.
 *
 * reads.bin: int64 n, int32 len[n], uint8 nt4 bases (concatenated)
 * jobs.bin:  int64 n, int8 mat[25], n x 56-byte mbg_dp_job_t, int64 nq, nq query bytes, int64 nt, nt target bytes
 *
 *   cpu_hotpath ll <jobs.bin> <n_threads>
 *
 * Mate-rescue local SW: ksw_ll_qinit + ksw_ll_u8_core / ksw_ll_i16_core (ksw2_ll_sse.c:49-333) per
 * job, as pe.c:268-272 calls them.  jobs.bin: int64 n, int8 mat[25], int32 gapo, gape, xtra,
 * n x 32-byte mbg_ll_job_t, int64 nq, nq query bytes, int64 nt, nt target bytes.
 *
 *   cpu_hotpath anchor <idx_prefix> <intv.bin> <max_occ> <n_threads>
 *
 * mb_anchor (seed.c:268-339, no methylation conversion) per read on given SMEM intervals, as
 * map-algo.c:610 calls it, then mb_lchain_dp (lchain.c:145-240) on its anchors as map-algo.c:616 does.
 * intv.bin: int64 n, int32 max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc, float pen_gap,
 * int32 qlen[n], int64 off[n+1], off[n] x 32-byte mb_sai_t.  Prints the wall time of the loop, the
 * thread-seconds spent in each of the two functions, and the anchor and chain counts.
 *
 *   cpu_hotpath bseq <reads.fq[.gz]> <chunk_size>
 *
 * The reference's reader: mb_bseq_read (bseq.c:85-131) in chunks of chunk_size bases, frag mode and
 * min_cnt 40000 as map-main.c:114-125 calls it for one interleaved file, until the input ends.
 * Prints the wall time, the records and the bases.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "mbpriv.h"
#include "bseq.h"
#include "ksw2.h"

typedef struct {
	int64_t qoff, toff;
	int32_t qlen, tlen;
	int8_t q, e, q2, e2;
	int32_t w, zdrop, end_bonus, flag;
	int64_t cigar_off;
} job_t;

static double now(void) { struct timeval tv; gettimeofday(&tv, 0); return tv.tv_sec + tv.tv_usec * 1e-6; }

typedef struct {
	mb_idx_t *idx;
	int64_t n_reads, *off;
	int32_t *len;
	uint8_t *seq;
	int64_t n_jobs;
	job_t *jobs;
	uint8_t *qb, *tb;
	int8_t mat[25];
	volatile int64_t next;
	int64_t n_anchor[256];
	int64_t n_cells[256];
} shared_t;

typedef struct { shared_t *s; int tid; } targ_t;

static void *seed_worker(void *a_)
{
	targ_t *a = (targ_t*)a_;
	shared_t *s = a->s;
	mb_sai_v v[48];
	mb_anchor_v av = {0, 0, 0};
	uint8_t *sq[48];
	int32_t ln[48];
	memset(v, 0, sizeof(v));
	for (;;) {
		int64_t b = __sync_fetch_and_add(&s->next, 48), e, i;
		if (b >= s->n_reads) break;
		e = b + 48 < s->n_reads ? b + 48 : s->n_reads;
		for (i = b; i < e; ++i) sq[i - b] = s->seq + s->off[i], ln[i - b] = s->len[i];
		mb_seed_intv_batch(0, s->idx->bwt, (int32_t)(e - b), ln, sq, 19, 10, v);
		for (i = b; i < e; ++i) {
			mb_anchor(0, s->idx, &v[i - b], 19, ln[i - b], sq[i - b], L2B_METH_NONE, 250, &av);
			s->n_anchor[a->tid] += av.n;
		}
	}
	for (int i = 0; i < 48; ++i) free(v[i].a);
	free(av.a);
	return 0;
}

static void *dp_worker(void *a_)
{
	targ_t *a = (targ_t*)a_;
	shared_t *s = a->s;
	ksw_extz_t ez;
	memset(&ez, 0, sizeof(ez));
	for (;;) {
		int64_t b = __sync_fetch_and_add(&s->next, 64), e, i;
		if (b >= s->n_jobs) break;
		e = b + 64 < s->n_jobs ? b + 64 : s->n_jobs;
		for (i = b; i < e; ++i) {
			const job_t *j = &s->jobs[i];
			if (j->q == j->q2 && j->e == j->e2)
				ksw_extz2_sse(0, j->qlen, s->qb + j->qoff, j->tlen, s->tb + j->toff, 5, s->mat, j->q, j->e, j->w, j->zdrop, j->end_bonus, j->flag, &ez);
			else
				ksw_extd2_sse(0, j->qlen, s->qb + j->qoff, j->tlen, s->tb + j->toff, 5, s->mat, j->q, j->e, j->q2, j->e2, j->w, j->zdrop, j->end_bonus, j->flag, &ez);
			s->n_cells[a->tid] += ez.max; /* consume the result */
		}
	}
	free(ez.cigar);
	return 0;
}

typedef struct { int64_t qoff, toff; int32_t qlen, tlen, size, pad; } ll_job_t;
typedef struct {
	int64_t n_jobs;
	ll_job_t *jobs;
	uint8_t *qb, *tb;
	int8_t mat[25];
	int32_t gapo, gape, xtra;
	volatile int64_t next;
	int64_t sum[256];
} ll_shared_t;
typedef struct { ll_shared_t *s; int tid; } ll_targ_t;

static void *ll_worker(void *a_)
{
	ll_targ_t *a = (ll_targ_t*)a_;
	ll_shared_t *s = a->s;
	for (;;) {
		int64_t b = __sync_fetch_and_add(&s->next, 64), e, i;
		if (b >= s->n_jobs) break;
		e = b + 64 < s->n_jobs ? b + 64 : s->n_jobs;
		for (i = b; i < e; ++i) {
			const ll_job_t *j = &s->jobs[i];
			void *qp = ksw_ll_qinit(0, j->size, j->qlen, s->qb + j->qoff, 5, s->mat);
			ksw_llrst_t r = j->size <= 1 ? ksw_ll_u8_core(qp, j->tlen, s->tb + j->toff, s->gapo, s->gape, s->xtra)
			                             : ksw_ll_i16_core(qp, j->tlen, s->tb + j->toff, s->gapo, s->gape, s->xtra);
			free(qp);
			s->sum[a->tid] += r.score + r.te + r.qe + r.score2; /* consume the result */
		}
	}
	return 0;
}

static void *slurp(const char *fn, size_t *len);

static int ll_main(const char *fn, int nt)
{
	ll_shared_t s;
	pthread_t th[256];
	ll_targ_t ta[256];
	size_t jl;
	uint8_t *p = (uint8_t*)slurp(fn, &jl);
	int64_t n, sum = 0;
	double t0, t;
	int i;
	if (!p) { fprintf(stderr, "cannot read %s\n", fn); return 1; }
	memset(&s, 0, sizeof(s));
	memcpy(&s.n_jobs, p, 8); p += 8;
	memcpy(s.mat, p, 25); p += 25;
	memcpy(&s.gapo, p, 4); memcpy(&s.gape, p + 4, 4); memcpy(&s.xtra, p + 8, 4); p += 12;
	s.jobs = (ll_job_t*)malloc(s.n_jobs * sizeof(ll_job_t) + 1);
	memcpy(s.jobs, p, s.n_jobs * sizeof(ll_job_t)); p += s.n_jobs * sizeof(ll_job_t);
	memcpy(&n, p, 8); p += 8; s.qb = p; p += n; memcpy(&n, p, 8); p += 8; s.tb = p;
	t0 = now();
	for (i = 0; i < nt; ++i) ta[i].s = &s, ta[i].tid = i, pthread_create(&th[i], 0, ll_worker, &ta[i]);
	for (i = 0; i < nt; ++i) pthread_join(th[i], 0);
	t = now() - t0;
	for (i = 0; i < nt; ++i) sum += s.sum[i];
	printf("{\"threads\": %d, \"n_jobs\": %ld, \"ll_s\": %.6f, \"checksum\": %ld}\n", nt, (long)s.n_jobs, t, (long)sum);
	return 0;
}

static void run(shared_t *s, int nt, void *(*fn)(void*))
{
	pthread_t th[256];
	targ_t ta[256];
	int i;
	s->next = 0;
	for (i = 0; i < nt; ++i) ta[i].s = s, ta[i].tid = i, pthread_create(&th[i], 0, fn, &ta[i]);
	for (i = 0; i < nt; ++i) pthread_join(th[i], 0);
}

static void *slurp(const char *fn, size_t *len)
{
	FILE *fp = fopen(fn, "rb");
	void *p;
	if (!fp) return 0;
	fseek(fp, 0, SEEK_END);
	*len = (size_t)ftell(fp);
	fseek(fp, 0, SEEK_SET);
	p = malloc(*len + 1);
	if (fread(p, 1, *len, fp) != *len) { fclose(fp); free(p); return 0; }
	fclose(fp);
	return p;
}

typedef struct {
	mb_idx_t *idx;
	int64_t n, *off;
	int32_t *qlen, max_occ;
	const mb_sai_t *u;
	int32_t cp[6];
	float pen_gap;
	volatile int64_t next;
	int64_t n_anchor[256], n_chain[256];
	double t_anchor[256], t_chain[256];
} anc_shared_t;

typedef struct { anc_shared_t *s; int tid; } anc_targ_t;

static void *anchor_worker(void *a_)
{
	anc_targ_t *a = (anc_targ_t*)a_;
	anc_shared_t *s = a->s;
	mb_sai_v v = {0, 0, 0};
	mb_anchor_v av = {0, 0, 0};
	for (;;) {
		int64_t b = __sync_fetch_and_add(&s->next, 64), e, i;
		if (b >= s->n) break;
		e = b + 64 < s->n ? b + 64 : s->n;
		for (i = b; i < e; ++i) {
			size_t k = (size_t)(s->off[i + 1] - s->off[i]);
			if (k > v.m) v.m = k, v.a = (mb_sai_t*)realloc(v.a, k * sizeof(mb_sai_t));
			memcpy(v.a, s->u + s->off[i], k * sizeof(mb_sai_t));     /* mb_anchor sorts them in place */
			v.n = k;
			if (k == 0) continue;                                      /* map-algo.c:595-598 */
			{
				double t0 = now(), t1;
				mb_anchor_t *b;
				uint64_t *w = 0;
				int n_u = 0;
				mb_anchor(0, s->idx, &v, 19, s->qlen[i], 0, L2B_METH_NONE, s->max_occ, &av);
				s->n_anchor[a->tid] += av.n;
				t1 = now();
				b = mb_lchain_dp(0, s->idx->l2b, s->cp[0], s->cp[1], s->cp[2], s->cp[3], s->cp[4], s->cp[5], s->pen_gap, av.n,
				                 av.a, &n_u, &w);                       /* frees av.a */
				av.a = 0, av.n = av.m = 0;
				s->t_chain[a->tid] += now() - t1, s->t_anchor[a->tid] += t1 - t0;
				s->n_chain[a->tid] += n_u;
				free(b); free(w);
			}
		}
	}
	free(v.a);
	free(av.a);
	return 0;
}

static int anchor_main(const char *prefix, const char *fn, int max_occ, int nt)
{
	anc_shared_t s;
	pthread_t th[256];
	anc_targ_t ta[256];
	size_t len;
	uint8_t *p = (uint8_t*)slurp(fn, &len);
	int64_t sum = 0, n_chain = 0;
	double t0, t, t_a = 0, t_c = 0;
	int i;
	if (!p) { fprintf(stderr, "cannot read %s\n", fn); return 1; }
	memset(&s, 0, sizeof(s));
	s.idx = mb_idx_load(prefix, 0);
	if (!s.idx) { fprintf(stderr, "cannot load index %s\n", prefix); return 1; }
	memcpy(&s.n, p, 8); p += 8;
	memcpy(s.cp, p, 24); p += 24;
	memcpy(&s.pen_gap, p, 4); p += 4;
	s.qlen = (int32_t*)p; p += 4 * s.n;
	s.off = (int64_t*)malloc((s.n + 1) * 8);
	memcpy(s.off, p, (s.n + 1) * 8); p += (s.n + 1) * 8;
	s.u = (const mb_sai_t*)p;
	s.max_occ = max_occ;
	t0 = now();
	for (i = 0; i < nt; ++i) ta[i].s = &s, ta[i].tid = i, pthread_create(&th[i], 0, anchor_worker, &ta[i]);
	for (i = 0; i < nt; ++i) pthread_join(th[i], 0);
	t = now() - t0;
	for (i = 0; i < nt; ++i) sum += s.n_anchor[i], n_chain += s.n_chain[i], t_a += s.t_anchor[i], t_c += s.t_chain[i];
	printf("{\"threads\": %d, \"n_reads\": %ld, \"n_intv\": %ld, \"wall_s\": %.6f, \"anchor_thread_s\": %.6f, "
	       "\"chain_thread_s\": %.6f, \"n_anchors\": %ld, \"n_chains\": %ld}\n", nt, (long)s.n, (long)s.off[s.n], t, t_a, t_c,
	       (long)sum, (long)n_chain);
	mb_idx_destroy(s.idx);
	return 0;
}

static int bseq_main(const char *fn, int64_t chunk)
{
	mb_bseq_file_t *fp = mb_bseq_open(fn);
	int64_t n_rec = 0, n_bases = 0;
	double t0;
	if (!fp) { fprintf(stderr, "cannot open %s\n", fn); return 1; }
	t0 = now();
	for (;;) {
		int n = 0, i;
		mb_bseq1_t *s = mb_bseq_read(fp, chunk, 1, 0, 1, 40000, chunk, &n);
		if (n == 0) { free(s); break; }
		for (i = 0; i < n; ++i) {
			n_bases += s[i].l_seq;
			free(s[i].name); free(s[i].seq); free(s[i].qual); free(s[i].comment);
		}
		n_rec += n;
		free(s);
	}
	printf("{\"bseq_s\": %.6f, \"records\": %ld, \"bases\": %ld}\n", now() - t0, (long)n_rec, (long)n_bases);
	mb_bseq_close(fp);
	return 0;
}

int main(int argc, char *argv[])
{
	shared_t s;
	size_t rl, jl;
	uint8_t *r, *jb, *p;
	int nt, i;
	double t0, t_seed, t_dp;
	int64_t n_anchor = 0;
	if (argc == 4 && strcmp(argv[1], "bseq") == 0) return bseq_main(argv[2], atoll(argv[3]));
	if (argc == 6 && strcmp(argv[1], "anchor") == 0) {
		nt = atoi(argv[5]);
		return anchor_main(argv[2], argv[3], atoi(argv[4]), nt < 1 ? 1 : nt > 256 ? 256 : nt);
	}
	if (argc == 4 && strcmp(argv[1], "ll") == 0) {
		nt = atoi(argv[3]);
		return ll_main(argv[2], nt < 1 ? 1 : nt > 256 ? 256 : nt);
	}
	if (argc < 5) { fprintf(stderr, "usage: cpu_hotpath <idx_prefix> <reads.bin> <jobs.bin> <n_threads>\n"); return 1; }
	memset(&s, 0, sizeof(s));
	nt = atoi(argv[4]);
	if (nt < 1) nt = 1;
	if (nt > 256) nt = 256;
	if (strcmp(argv[1], "-") != 0) {
		s.idx = mb_idx_load(argv[1], 0);
		if (!s.idx) { fprintf(stderr, "cannot load index %s\n", argv[1]); return 1; }
	}
	r = (uint8_t*)slurp(argv[2], &rl);
	jb = (uint8_t*)slurp(argv[3], &jl);
	if (!r || !jb) { fprintf(stderr, "cannot read inputs\n"); return 1; }
	memcpy(&s.n_reads, r, 8);
	s.len = (int32_t*)(r + 8);
	s.seq = r + 8 + 4 * s.n_reads;
	s.off = (int64_t*)malloc((s.n_reads + 1) * 8);
	s.off[0] = 0;
	for (i = 0; i < s.n_reads; ++i) s.off[i + 1] = s.off[i] + s.len[i];
	p = jb;
	memcpy(&s.n_jobs, p, 8); p += 8;
	memcpy(s.mat, p, 25); p += 25;
	s.jobs = (job_t*)malloc(s.n_jobs * sizeof(job_t) + 1);
	memcpy(s.jobs, p, s.n_jobs * sizeof(job_t)); p += s.n_jobs * sizeof(job_t);
	{ int64_t n; memcpy(&n, p, 8); p += 8; s.qb = p; p += n; memcpy(&n, p, 8); p += 8; s.tb = p; }
	t0 = now();
	if (s.idx) run(&s, nt, seed_worker);
	t_seed = now() - t0;
	t0 = now();
	run(&s, nt, dp_worker);
	t_dp = now() - t0;
	for (i = 0; i < nt; ++i) n_anchor += s.n_anchor[i];
	printf("{\"threads\": %d, \"n_reads\": %ld, \"n_jobs\": %ld, \"n_anchors\": %ld, \"seed_anchor_s\": %.6f, \"dp_s\": %.6f}\n",
	       nt, (long)s.n_reads, (long)s.n_jobs, (long)n_anchor, t_seed, t_dp);
	if (s.idx) mb_idx_destroy(s.idx);
	return 0;
}
