/*
 * ref_pe_shim.c -- exposes the reference's mate-rescue steps, which pe.c keeps static, to the tests:
 * mb_ungap (pe.c:209-247) and mb_matesw_align (pe.c:249-312).  TEST INFRASTRUCTURE ONLY.
 * oracle/Makefile compiles it with pe.c included from where it lies in the reference tree (nothing
 * of it is copied here) into oracle/_ref/libmbref_pe.so.
 */
#include "pe.c"

/* mb_ungap on one query/window pair; out = {max_i, n_good, n_kmer}; returns max */
int32_t mbx_ungap(int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq, int32_t kmer, int32_t mt, int32_t out[3])
{
	return mb_ungap(0, qlen, qseq, tlen, tseq, kmer, (l2b_meth_t)mt, &out[0], &out[1], &out[2]);
}

/* the fields of a hit the tests compare, after mb_update_extra */
typedef struct {
	int32_t rescued, qs, qe, score, score0, subsc, dp_max, dp_max2, dp_score, dp_max0, mlen, blen, n_ambi, n_cigar;
	int64_t ts, te;
} mbx_hit_t;

static void mbx_opt(mb_opt_t *o, const int32_t p[11])
{
	mb_opt_init(o);
	o->a = p[0], o->b = p[1], o->b_ts = p[2], o->b_ambi = p[3], o->q = p[4], o->e = p[5];
	o->bw = p[6], o->zdrop = p[7], o->end_bonus = p[8], o->min_dp_max = p[9], o->min_len = p[10];
}

static void mbx_out(const mb_hit_t *h, mbx_hit_t *r, uint32_t *cigar, int32_t cap)
{
	memset(r, 0, sizeof(*r));
	r->rescued = h->rescued, r->qs = h->qs, r->qe = h->qe, r->ts = h->ts, r->te = h->te;
	r->score = h->score, r->score0 = h->score0, r->subsc = h->subsc, r->mlen = h->mlen, r->blen = h->blen;
	if (h->p) {
		r->dp_max = h->p->dp_max, r->dp_max2 = h->p->dp_max2, r->dp_score = h->p->dp_score, r->dp_max0 = h->p->dp_max0;
		r->n_ambi = h->p->n_ambi, r->n_cigar = h->p->n_cigar;
		memcpy(cigar, h->p->cigar, 4 * (size_t)(h->p->n_cigar < cap ? h->p->n_cigar : cap));
	}
}

/* mb_matesw_align on one window: p = {a, b, b_ts, b_ambi, q, e, bw, zdrop, end_bonus, min_dp_max, min_len};
 * ez_out = {max, zdropped, max_q, max_t, mqe, mqe_t, mte, mte_q, score, reach_end, n_cigar} of the
 * extension DP (all 0 when it did not run) */
void mbx_matesw_align(const int32_t p[11], int32_t qlen, uint8_t *qseq, int32_t tlen, uint8_t *tseq, int32_t min_sc, int32_t mt,
                      mbx_hit_t *r, uint32_t *cigar, int32_t cap, int32_t ez_out[11], uint32_t *ez_cigar, int32_t ez_cap)
{
	mb_opt_t opt;
	mb_hit_t h;
	ksw_extz_t ez;
	mbx_opt(&opt, p);
	memset(&ez, 0, sizeof(ez));
	mb_matesw_align(0, &opt, qlen, qseq, tlen, tseq, &h, min_sc, (l2b_meth_t)mt, &ez);
	mbx_out(&h, r, cigar, cap);
	ez_out[0] = ez.max, ez_out[1] = ez.zdropped, ez_out[2] = ez.max_q, ez_out[3] = ez.max_t, ez_out[4] = ez.mqe;
	ez_out[5] = ez.mqe_t, ez_out[6] = ez.mte, ez_out[7] = ez.mte_q, ez_out[8] = ez.score, ez_out[9] = ez.reach_end;
	ez_out[10] = ez.n_cigar;
	if (ez.n_cigar > 0) memcpy(ez_cigar, ez.cigar, 4 * (size_t)(ez.n_cigar < ez_cap ? ez.n_cigar : ez_cap));
	free(ez.cigar);
	free(h.p);
}

/* the host's share of a rescue whose DP ran elsewhere: the hit built from pre-update fields as
 * pe.c:293-300 fills it, then the reference's mb_append_cigar and mb_update_extra (pe.c:294,307) */
void mbx_finish(const int32_t p[11], int32_t mt, const uint8_t *qseq, const uint8_t *tseq, int32_t qs, int32_t qe, int64_t ts,
                int64_t te, int32_t ez_max, int32_t score, int32_t score2, int32_t dp_max2, int32_t n_cigar, const uint32_t *cigar0,
                mbx_hit_t *r, uint32_t *cigar, int32_t cap)
{
	mb_opt_t opt;
	mb_hit_t h;
	int8_t mat[25];
	mbx_opt(&opt, p);
	memset(&h, 0, sizeof(h));
	mb_append_cigar(&h, n_cigar, cigar0);
	h.rescued = 1, h.qe = qe, h.te = te, h.ts = ts, h.qs = qs;
	h.p->dp_max = h.p->dp_score = ez_max, h.p->dp_max2 = dp_max2;
	h.score = h.score0 = score, h.subsc = score2;
	h.cnt = 0, h.as = -1, h.parent = MB_PARENT_UNSET;
	ksw_gen_nt4_mat(mat, opt.a, opt.b, opt.b_ts, opt.b_ambi, mt);
	mb_update_extra(0, &h, &qseq[h.qs], &tseq[h.ts], mat, opt.q, opt.e, opt.flag, 0);
	mbx_out(&h, r, cigar, cap);
	free(h.p);
}

/* the reference's mb_update_extra (align.c:219-283, with mb_fix_cigar) on one hit, no cs/MD/=X
 * output: qseq and tseq point at the hit's first bases, as its callers pass them */
void mbx_update_extra(int32_t qs, int32_t qe, int64_t ts, int64_t te, int32_t rev, int32_t n_cigar, const uint32_t *cigar0,
                      const uint8_t *qseq, const uint8_t *tseq, const int8_t *mat, int8_t q, int8_t e, int32_t log_gap,
                      mbx_hit_t *r, uint32_t *cigar, int32_t cap)
{
	mb_hit_t h;
	memset(&h, 0, sizeof(h));
	mb_append_cigar(&h, n_cigar, cigar0);
	h.qs = qs, h.qe = qe, h.ts = ts, h.te = te, h.rev = rev;
	mb_update_extra(0, &h, qseq, tseq, mat, q, e, 0, log_gap);
	mbx_out(&h, r, cigar, cap);
	free(h.p);
}

/* the reference's mb_anchor (seed.c:268-339) on one read's intervals, no methylation conversion;
 * returns the number of anchors (at most cap copied to out) */
int64_t mbx_anchor(const mb_idx_t *idx, const mb_sai_t *u0, int64_t n_u, int32_t qlen, int32_t max_occ, mb_anchor_t *out, int64_t cap)
{
	mb_sai_v u;
	mb_anchor_v v = {0, 0, 0};
	int64_t n;
	u.n = u.m = (size_t)n_u;
	u.a = (mb_sai_t*)malloc(sizeof(mb_sai_t) * (n_u + 1));
	memcpy(u.a, u0, sizeof(mb_sai_t) * n_u);
	mb_anchor(0, idx, &u, 19, qlen, 0, L2B_METH_NONE, max_occ, &v);
	n = v.n;
	memcpy(out, v.a, sizeof(mb_anchor_t) * (size_t)(n < cap ? n : cap));
	free(u.a); free(v.a);
	return n;
}

/* mb_lchain_dp (lchain.c:145-240) on a copy of a[0..n) (the reference frees its input): the chains
 * (score<<32 | count) to u_out and their anchors to b_out; returns n_u, *n_b the anchor count. */
int32_t mbx_lchain(const mb_idx_t *idx, int32_t max_dist_x, int32_t max_dist_y, int32_t bw, int32_t max_skip, int32_t max_iter,
                   int32_t min_sc, float pen_gap, int64_t n, const mb_anchor_t *a, uint64_t *u_out, mb_anchor_t *b_out,
                   int64_t *n_b)
{
	mb_anchor_t *a2 = (mb_anchor_t*)malloc(sizeof(mb_anchor_t) * (n + 1)), *b;
	uint64_t *u = 0;
	int32_t n_u = 0, i;
	int64_t k = 0;
	memcpy(a2, a, sizeof(mb_anchor_t) * n);
	b = mb_lchain_dp(0, idx->l2b, max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc, pen_gap, n, a2, &n_u, &u);
	for (i = 0; i < n_u; ++i) u_out[i] = u[i], k += (int32_t)u[i];
	if (b) memcpy(b_out, b, sizeof(mb_anchor_t) * k);
	*n_b = k;
	free(b); free(u);
	return n_u;
}
