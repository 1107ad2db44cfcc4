/*
 * cpu_rescue.c -- times the reference's mate-rescue step on the host cores (the CPU baseline of
 * bench.py --rescue-batch).  TEST/BASELINE INFRASTRUCTURE: linked against oracle/_ref/libmbref_pe.so
 * (the reference compiled where it lies, with mb_ungap and mb_matesw_align exported by
 * ref_pe_shim.c), never part of libmbg.
 *
 *   cpu_rescue <cands.bin> <n_threads>
 *
 * Per candidate, as mb_matesw_core runs it after l2b_getseq (pe.c:342-349): mb_ungap over the
 * window, the narrowing to a clear diagonal, and mb_matesw_align (ksw_ll, ksw_extz2 of the reversed
 * prefixes, mb_append_cigar + mb_update_extra) when the vote passes.  Work-stealing over threads.
 * cands.bin: int32 p[11] (a, b, b_ts, b_ambi, q, e, bw, zdrop, end_bonus, min_dp_max, min_len),
 * int32 mt, int64 n, n x {int64 qoff, int64 woff, int32 qlen, int32 wlen, int32 min_sc, int32 pad},
 * int64 nq, nq read bytes, int64 nw, nw window bytes.
 * Prints one JSON line: wall time and the number of rescued candidates.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

typedef struct {
	int32_t rescued, qs, qe, score, score0, subsc, dp_max, dp_max2, dp_score, dp_max0, mlen, blen, n_ambi, n_cigar;
	int64_t ts, te;
} mbx_hit_t;
int32_t mbx_ungap(int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq, int32_t kmer, int32_t mt, int32_t out[3]);
void mbx_matesw_align(const int32_t p[11], int32_t qlen, uint8_t *qseq, int32_t tlen, uint8_t *tseq, int32_t min_sc, int32_t mt,
                      mbx_hit_t *r, uint32_t *cigar, int32_t cap, int32_t ez_out[11], uint32_t *ez_cigar, int32_t ez_cap);

typedef struct { int64_t qoff, woff; int32_t qlen, wlen, min_sc, pad; } cand_t;

typedef struct {
	int32_t p[11], mt;
	int64_t n;
	cand_t *c;
	uint8_t *qb, *wb;
	volatile int64_t next;
	int64_t rescued[256];
} shared_t;
typedef struct { shared_t *s; int tid; } targ_t;

static double now(void) { struct timeval tv; gettimeofday(&tv, 0); return tv.tv_sec + tv.tv_usec * 1e-6; }

static void *worker(void *a_)
{
	targ_t *a = (targ_t*)a_;
	shared_t *s = a->s;
	uint8_t *q = (uint8_t*)malloc(1 << 16), *w = (uint8_t*)malloc(1 << 20);
	uint32_t *cig = (uint32_t*)malloc(4 << 20), *ezc = (uint32_t*)malloc(4 << 20);
	for (;;) {
		int64_t b = __sync_fetch_and_add(&s->next, 64), e, i;
		if (b >= s->n) break;
		e = b + 64 < s->n ? b + 64 : s->n;
		for (i = b; i < e; ++i) {
			const cand_t *c = &s->c[i];
			int32_t ug[3], max_ug, ts2 = 0, te2 = c->wlen, ez[11];
			mbx_hit_t h;
			if (c->wlen <= c->qlen) continue;
			memcpy(q, s->qb + c->qoff, c->qlen);           /* mb_matesw_align reverses prefixes in place */
			memcpy(w, s->wb + c->woff, c->wlen);
			max_ug = mbx_ungap(c->qlen, q, c->wlen, w, 7, s->mt, ug);
			if (max_ug >= 10 && max_ug >= c->qlen >> 1 && ug[1] == 1) {   /* pe.c:343-348 */
				ts2 = ug[0] - c->qlen / 2, te2 = ts2 + c->qlen * 2;
				if (ts2 < 0) ts2 = 0;
				if (te2 > c->wlen) te2 = c->wlen;
			}
			if (max_ug >= 10 || max_ug >= ug[2] * 0.33) {
				mbx_matesw_align(s->p, c->qlen, q, te2 - ts2, w + ts2, c->min_sc, s->mt, &h, cig, 1 << 20, ez, ezc, 1 << 20);
				s->rescued[a->tid] += h.rescued;
			}
		}
	}
	free(q); free(w); free(cig); free(ezc);
	return 0;
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

int main(int argc, char *argv[])
{
	shared_t s;
	pthread_t th[256];
	targ_t ta[256];
	size_t len;
	uint8_t *p;
	int64_t nb, rescued = 0;
	int nt, i;
	double t0, t;
	if (argc < 3) { fprintf(stderr, "usage: cpu_rescue <cands.bin> <n_threads>\n"); return 1; }
	memset(&s, 0, sizeof(s));
	nt = atoi(argv[2]);
	nt = nt < 1 ? 1 : nt > 256 ? 256 : nt;
	if ((p = (uint8_t*)slurp(argv[1], &len)) == 0) { fprintf(stderr, "cannot read %s\n", argv[1]); return 1; }
	memcpy(s.p, p, 44); p += 44;
	memcpy(&s.mt, p, 4); p += 4;
	memcpy(&s.n, p, 8); p += 8;
	s.c = (cand_t*)p; p += s.n * sizeof(cand_t);
	memcpy(&nb, p, 8); p += 8; s.qb = p; p += nb;
	memcpy(&nb, p, 8); p += 8; s.wb = p;
	t0 = now();
	for (i = 0; i < nt; ++i) ta[i].s = &s, ta[i].tid = i, pthread_create(&th[i], 0, worker, &ta[i]);
	for (i = 0; i < nt; ++i) pthread_join(th[i], 0);
	t = now() - t0;
	for (i = 0; i < nt; ++i) rescued += s.rescued[i];
	printf("{\"threads\": %d, \"n\": %ld, \"rescued\": %ld, \"rescue_s\": %.6f}\n", nt, (long)s.n, (long)rescued, t);
	return 0;
}
