/*
 * mb_oracle.c -- CPU restatement of minibwa 0.7-r421's seeding + extension hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Linked by tests/, smoke() and
 * bench.py's cpu_baseline leg; the product never links it.
 *
 * What is restated, and what pins it (each pin is a test in tests/test_oracle.py):
 *   FM-index load/rank/extend  bwt.c:176-250,648-683   rank: brute-force count over the decoded
 *                                                      BWT; the `bench -b 2a` checksum on chrM
 *   SA locate                  bwt.c:469-541           brute-force suffix sort of tiny texts;
 *                                                      `bench -b sa` / `-b msa` checksums on chrM
 *   10-mer cache               bwt.c:552-591           equals stepwise extension of every 10-mer
 *   SMEM + two-pass seeding    bwt.c:253-305, seed.c:24-54
 *                                                      brute-force (l,c)-SMEM definition on tiny
 *                                                      references; the compiled reference
 *   ksw_extd2 / ksw_extz2      ksw2_extd2_sse.c:27-403, ksw2_extz2_sse.c:16-296, ksw2.h:124-197
 *                                                      textbook Gotoh global score; the compiled
 *                                                      reference on fuzzed jobs (all flag variants)
 *   ksw_ll (u8 / i16 cores)    ksw2_ll_sse.c:49-333     textbook Gotoh local score (brute force on
 *                                                      small jobs); the compiled reference on fuzzed
 *                                                      jobs (both sizes, SUBO/STOP, saturation)
 *   mate rescue                pe.c:209-312,342-349     mb_ungap: brute-force vote count; mb_ungap and
 *     (ungap, matesw_align)                            mb_matesw_align of the compiled reference
 *   mb_anchor (no methylation) seed.c:109-339          the compiled reference's mb_anchor on chrM
 *                                                      and repeat-rich indexes (klib sort ties)
 *   mb_fix_cigar + mb_update_extra  align.c:70-258      the compiled reference's mb_update_extra on
 *                                                      hits from the DP (both gap models); the
 *                                                      score of the fixed CIGAR (invariant)
 *
 * The DP restatement is scalar but keeps the SSE code's observable semantics: int8 wrap-around
 * arithmetic, the 16-cell-aligned band (cells outside the exact band are computed and kept), the
 * byte layout of the scratch block (stale scores, reads past the target into the reversed query),
 * the 4-lane argmax tie-break of the exact-max pass, and the rotated backtrack matrix.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mb_oracle.h"

#define M56 ((1ULL << 56) - 1)

/*******************
 * FM-index access *
 *******************/

static uint64_t orc_data_len(uint64_t len) /* bwt.c:52-58 */
{
	uint64_t nb = (len + 127) / 128;
	return nb * 4 + (nb + 1) * 4;
}

orc_bwt_t *orc_bwt_load(const char *fn) /* bwt.c:648-683 */
{
	FILE *fp;
	char magic[4];
	uint64_t hdr[5];
	orc_bwt_t *b;
	if ((fp = fopen(fn, "rb")) == 0) return 0;
	if (fread(magic, 1, 4, fp) != 4 || memcmp(magic, "MBW\2", 4) != 0) { fclose(fp); return 0; }
	b = (orc_bwt_t*)calloc(1, sizeof(*b));
	if (fread(&b->sa_bit, 4, 1, fp) != 1 || fread(hdr, 8, 5, fp) != 5) goto fail;
	b->primary = hdr[0];
	b->L2[0] = 0;
	memcpy(&b->L2[1], &hdr[1], 32);
	b->seq_len = b->L2[4];
	b->data_len = orc_data_len(b->seq_len);
	b->data = (uint64_t*)malloc(b->data_len * 8);
	if (fread(b->data, 8, b->data_len, fp) != b->data_len) goto fail;
	if (fread(&b->n_sa, 8, 1, fp) != 1) goto fail;
	if (b->sa_bit != (uint32_t)-1 && b->n_sa > 0) {
		if (b->n_sa != ((b->seq_len + (1ULL << b->sa_bit)) >> b->sa_bit)) goto fail;
		b->sa = (uint64_t*)malloc(b->n_sa * 8);
		if (fread(b->sa, 8, b->n_sa, fp) != b->n_sa) goto fail;
	}
	fclose(fp);
	return b;
fail:
	fclose(fp);
	orc_bwt_destroy(b);
	return 0;
}

void orc_bwt_destroy(orc_bwt_t *b)
{
	if (b == 0) return;
	free(b->data); free(b->sa); free(b->pre); free(b);
}

/* number of symbol c among the first n (0..32) 2-bit symbols of w */
static inline uint64_t sym_count(uint64_t w, int c, int n)
{
	static const uint64_t pat[4] = { 0, 0x5555555555555555ULL, 0xAAAAAAAAAAAAAAAAULL, ~0ULL };
	uint64_t e = ~(w ^ pat[c]);
	e = e & (e >> 1) & 0x5555555555555555ULL;
	if (n < 32) e &= (1ULL << (2 * n)) - 1;
	return (uint64_t)__builtin_popcountll(e);
}

/* occurrences of each base among $-removed BWT symbols [0, j] (inclusive) -- bwt.c:160-193 */
static void occ_incl(const orc_bwt_t *b, uint64_t j, uint64_t cnt[4])
{
	const uint64_t *p = b->data + (j >> 7 << 3);
	int c, in = (int)(j & 127), half = in >= 64, n = (in & 63) + 1;
	for (c = 0; c < 4; ++c) {
		uint64_t v = (p[c] & M56) + (half ? p[c] >> 56 : 0);
		const uint64_t *w = p + 4 + (half ? 2 : 0);
		v += n > 32 ? sym_count(w[0], c, 32) + sym_count(w[1], c, n - 32) : sym_count(w[0], c, n);
		cnt[c] = v;
	}
}

void orc_rank1a(const orc_bwt_t *b, uint64_t k, uint64_t cnt[4]) /* bwt.c:176-193 */
{
	if (k == 0) { memset(cnt, 0, 32); return; }
	--k;
	k -= (k >= b->primary);
	occ_incl(b, k, cnt);
}

static inline int bwt_b0(const orc_bwt_t *b, uint64_t k) /* bwt.c:119 */
{
	return b->data[(k >> 7 << 3) + 4 + ((k & 127) >> 5)] >> ((k & 31) << 1) & 3;
}

/* counts of each base among positions [0, o] of the 16-symbol group g (0..7) of block p */
static void grp_count(const uint64_t *p, int g, int o, uint64_t cnt[4])
{
	uint64_t w = p[4 + (g >> 1)] >> ((g & 1) * 32);
	int c;
	w &= 0xffffffffULL;
	for (c = 0; c < 4; ++c) cnt[c] += sym_count(w, c, o + 1);
}

void orc_rank2a(const orc_bwt_t *b, uint64_t k, uint64_t l, uint64_t ck[4], uint64_t cl[4]) /* bwt.c:195-231 */
{
	uint64_t k1 = k - 1, l1 = l - 1;
	k1 -= (k1 >= b->primary);
	l1 -= (l1 >= b->primary);
	if (k1 >> 7 != l1 >> 7 || k == 0 || l == 0) {
		orc_rank1a(b, k, ck);
		orc_rank1a(b, l, cl);
	} else if (l - k == 1) {
		/* bwt.c:205-209: adds the symbol at $-removed index k-(k>primary); at k==primary that row
		 * holds '$', so this differs from rank1a(l) */
		orc_rank1a(b, k, ck);
		memcpy(cl, ck, 32);
		++cl[bwt_b0(b, k - (k > b->primary))];
	} else {
		/* bwt.c:210-230, one block.  Counts advance in 16-symbol groups from the group of k1; for
		 * l1 < k1 (never on the mapping path, which ranks x and x+size) the l1 count is taken inside
		 * k1's group, as the reference does. */
		const uint64_t *p = b->data + (k1 >> 7 << 3);
		int c, half = (k1 & 127) >= 64, g, gk = (int)((k1 & 127) >> 4), gl = (int)((l1 & 127) >> 4);
		uint64_t base[4], x[4] = {0, 0, 0, 0}, y[4];
		for (c = 0; c < 4; ++c) base[c] = (p[c] & M56) + (half ? p[c] >> 56 : 0);
		for (g = half ? 4 : 0; g < gk; ++g) grp_count(p, g, 15, x);
		memcpy(y, x, 32);
		grp_count(p, gk, (int)(k1 & 15), x);
		for (g = gk; g < gl; ++g) grp_count(p, g, 15, y);
		grp_count(p, g, (int)(l1 & 15), y);
		for (c = 0; c < 4; ++c) ck[c] = base[c] + x[c], cl[c] = base[c] + y[c];
	}
}

void orc_extend(const orc_bwt_t *b, const orc_sai_t *ik, orc_sai_t ok[4], int is_back) /* bwt.c:237-250 */
{
	uint64_t tk[4], tl[4];
	int i, f = !is_back, g = is_back;
	orc_rank2a(b, ik->x[f], ik->x[f] + ik->size, tk, tl);
	for (i = 0; i < 4; ++i) {
		ok[i].x[f] = b->L2[i] + 1 + tk[i];
		ok[i].size = tl[i] - tk[i];
		ok[i].info = 0;
	}
	ok[3].x[g] = ik->x[g] + (ik->x[f] <= b->primary && ik->x[f] + ik->size > b->primary);
	ok[2].x[g] = ok[3].x[g] + ok[3].size;
	ok[1].x[g] = ok[2].x[g] + ok[2].size;
	ok[0].x[g] = ok[1].x[g] + ok[1].size;
}

static void set_intv(const orc_bwt_t *b, int c, orc_sai_t *ik) /* bwt.h:75-81 */
{
	ik->x[0] = b->L2[c] + 1;
	ik->x[1] = b->L2[3 - c] + 1;
	ik->size = b->L2[c + 1] - b->L2[c];
	ik->info = 0;
}

void orc_bwt_cache(orc_bwt_t *b, int32_t len) /* bwt.c:552-591: interval of every len-mer */
{
	uint64_t z, n = 1ULL << (2 * len);
	free(b->pre);
	b->pre_len = len;
	b->pre = (orc_sai_t*)calloc(n, sizeof(orc_sai_t));
	for (z = 0; z < n; ++z) { /* key: last base in the top two bits, first base in the lowest */
		orc_sai_t p, ok[4];
		int d;
		set_intv(b, (int)(z >> (2 * (len - 1)) & 3), &p);
		for (d = len - 2; d >= 0; --d) {
			orc_extend(b, &p, ok, 1);
			p = ok[z >> (2 * d) & 3];
		}
		b->pre[z] = p;
	}
}

/********
 * SMEM *
 ********/

static int64_t smem_back(const orc_bwt_t *f, const uint8_t *q, int64_t st, int64_t pos, int64_t min_occ, orc_sai_t *p) /* bwt.c:253-277 */
{
	int64_t i = pos - 1;
	orc_sai_t ok[4];
	if (f->pre && pos - st >= f->pre_len) {
		uint64_t z = 0, l = 0;
		for (i = pos; l < (uint64_t)f->pre_len; --i, ++l)
			z = z << 2 | q[i];
		*p = f->pre[z];
	} else p->size = 0;
	if (p->size < (uint64_t)min_occ) {
		set_intv(f, q[pos], p);
		i = pos - 1;
	}
	for (; i >= st; --i) {
		int c = q[i];
		if (c > 3) break;
		orc_extend(f, p, ok, 1);
		if (ok[c].size < (uint64_t)min_occ) break;
		*p = ok[c];
	}
	return i;
}

int64_t orc_smem(const orc_bwt_t *f, uint32_t len, const uint8_t *q, int64_t x, int64_t min_len, int64_t min_occ, orc_sai_t *p) /* bwt.c:280-305 */
{
	int64_t i, j, xn;
	orc_sai_t ik, ok[4];
	p->size = ik.size = 0;
	if ((int64_t)len - x < min_len) return len;
	for (i = x, xn = -1; i < x + min_len; ++i)
		if (q[i] > 3) xn = i;
	if (xn >= 0) return xn + 1;
	i = smem_back(f, q, x, x + min_len - 1, min_occ, &ik);
	if (i >= x) return i + 1;
	for (j = x + min_len; j < (int64_t)len; ++j) {
		int c = 3 - q[j];
		if (q[j] > 3) break;
		orc_extend(f, &ik, ok, 0);
		if (ok[c].size < (uint64_t)min_occ) break;
		ik = ok[c];
	}
	*p = ik;
	p->info = (uint64_t)x << 32 | (uint64_t)j;
	if (j == (int64_t)len) return len;
	i = q[j] > 3 ? j : smem_back(f, q, x + 1, j, min_occ, &ik);
	return i + 1;
}

/* seed.c:24-54; returns the number of intervals (written up to cap) */
int64_t orc_seed_intv(const orc_bwt_t *b, int32_t len, const uint8_t *seq, int32_t min_len, int32_t max_sub_occ, orc_sai_t *out, int64_t cap)
{
	int64_t x = 0, n = 0, n0, i;
	orc_sai_t p;
	if (len <= 0) return 0;
	do {
		x = orc_smem(b, len, seq, x, min_len, 1, &p);
		if (p.size > 0) { if (n < cap) out[n] = p; ++n; }
	} while (x < len);
	n0 = n < cap ? n : cap;
	for (i = 0; i < n0; ++i) {
		uint32_t st = out[i].info >> 32, en = (uint32_t)out[i].info;
		int32_t sub;
		if (en - st < (uint32_t)min_len * 2 || out[i].size > (uint64_t)max_sub_occ) continue;
		x = st;
		sub = (int32_t)(en - st) / 2 > min_len ? (int32_t)(en - st) / 2 : min_len;
		do {
			x = orc_smem(b, en, seq, x, sub, out[i].size + 1, &p);
			if (p.size > out[i].size) { if (n < cap) out[n] = p; ++n; }
		} while (x < en);
	}
	return n;
}

/******
 * SA *
 ******/

static inline uint64_t inv_psi(const orc_bwt_t *b, uint64_t k) /* bwt.c:469-475 */
{
	uint64_t x = k - (k > b->primary), cnt[4];
	int c;
	if (k == b->primary) return 0;
	c = bwt_b0(b, x);
	orc_rank1a(b, k, cnt);
	return b->L2[c] + 1 + cnt[c];
}

uint64_t orc_sa(const orc_bwt_t *b, uint64_t k) /* bwt.c:499-509 */
{
	uint64_t s = 0, mask = (1ULL << b->sa_bit) - 1;
	while (k & mask) { ++s; k = inv_psi(b, k); }
	return s + b->sa[k >> b->sa_bit];
}

void orc_sa_batch(const orc_bwt_t *b, int64_t n, uint64_t *x) /* bwt.c:511-541: same values, in place */
{
	int64_t i;
	for (i = 0; i < n; ++i) x[i] = orc_sa(b, x[i]);
}

/*************
 * ksw2 DP   *
 *************/

#define NEG_INF (-0x40000000)

typedef struct { int n, cap, overflow; uint32_t *a; } cig_t;

static void cig_push(cig_t *c, uint32_t op, int len) /* ksw2.h:124-134 */
{
	if (c->n > 0 && (c->a[c->n - 1] & 0xf) == op) { c->a[c->n - 1] += (uint32_t)len << 4; return; }
	if (c->n < c->cap) c->a[c->n] = (uint32_t)len << 4 | op;
	else c->overflow = 1;
	c->n++;
}

/* ksw2.h:140-172 with is_rot=1, min_intron_len=0 */
static void backtrack(cig_t *c, int is_rev, const uint8_t *p, const int *off, const int *off_end, size_t n_col, int i0, int j0)
{
	int i = i0, j = j0, state = 0, k;
	while (i >= 0 && j >= 0) {
		int r = i + j, force = -1;
		uint32_t tmp;
		if (i < off[r]) force = 2;
		if (i > off_end[r]) force = 1;
		tmp = force < 0 ? p[(size_t)r * n_col + i - off[r]] : 0;
		if (state == 0) state = tmp & 7;
		else if (!(tmp >> (state + 2) & 1)) state = 0;
		if (state == 0) state = tmp & 7;
		if (force >= 0) state = force;
		if (state == 0) cig_push(c, 0, 1), --i, --j;
		else if (state == 1 || state == 3) cig_push(c, 2, 1), --i;
		else cig_push(c, 1, 1), --j;
	}
	if (i >= 0) cig_push(c, 2, i + 1);
	if (j >= 0) cig_push(c, 1, j + 1);
	if (!is_rev && c->n <= c->cap)
		for (k = 0; k < c->n >> 1; ++k) {
			uint32_t t = c->a[k]; c->a[k] = c->a[c->n - 1 - k]; c->a[c->n - 1 - k] = t;
		}
}

static void reset_ez(orc_ez_t *ez) /* ksw2.h:174-179 */
{
	ez->max_q = ez->max_t = ez->mqe_t = ez->mte_q = -1;
	ez->max = 0, ez->score = ez->mqe = ez->mte = NEG_INF;
	ez->n_cigar = 0, ez->zdropped = 0, ez->reach_end = 0;
}

static int apply_zdrop(orc_ez_t *ez, int32_t H, int r, int t, int zdrop, int e) /* ksw2.h:181-197, is_rot=1 */
{
	if (H > ez->max) {
		ez->max = H & 0x7fffffff, ez->max_t = t, ez->max_q = r - t;
	} else if (t >= ez->max_t && r - t >= ez->max_q) {
		int tl = t - ez->max_t, ql = (r - t) - ez->max_q, l = tl > ql ? tl - ql : ql - tl;
		if (zdrop >= 0 && ez->max - H > zdrop + l * e) { ez->zdropped = 1; return 1; }
	}
	return 0;
}

static inline int8_t i8max(int8_t a, int8_t b) { return a > b ? a : b; }
static inline int8_t i8min(int8_t a, int8_t b) { return a < b ? a : b; }
static inline uint8_t u8max(uint8_t a, uint8_t b) { return a > b ? a : b; }
static inline uint8_t u8min(uint8_t a, uint8_t b) { return a < b ? a : b; }

/* exact-max pass of ksw2_extd2_sse.c:318-351 / ksw2_extz2_sse.c:219-253: the argmax follows the
 * 4-lane SSE reduction (H[en0] seeded first; lanes t=st0+i mod 4 keep their first strict max;
 * lanes merged in order 0..3; then the scalar tail [en1,en0)). `sub` is subtracted per cell. */
static void exact_max(int32_t *H, const uint8_t *vb, const uint8_t *ub, int is_signed, int st0, int en0, int sub, int32_t *max_H_, int32_t *max_t_)
{
	int32_t max_H, max_t, HH[4], tt[4], t, i, en1 = st0 + (en0 - st0) / 4 * 4;
#define VAL(arr, k) (is_signed ? (int32_t)(int8_t)(arr)[k] : (int32_t)(arr)[k])
	max_H = H[en0] = en0 > 0 ? H[en0 - 1] + VAL(ub, en0) - sub : H[en0] + VAL(vb, en0) - sub;
	max_t = en0;
	for (i = 0; i < 4; ++i) HH[i] = max_H, tt[i] = max_t;
	for (t = st0; t < en1; t += 4)
		for (i = 0; i < 4; ++i) {
			H[t + i] += VAL(vb, t + i) - sub;
			if (H[t + i] > HH[i]) HH[i] = H[t + i], tt[i] = t;
		}
	for (i = 0; i < 4; ++i)
		if (max_H < HH[i]) max_H = HH[i], max_t = tt[i] + i;
	for (; t < en0; ++t) {
		H[t] += VAL(vb, t) - sub;
		if (H[t] > max_H) max_H = H[t], max_t = t;
	}
#undef VAL
	*max_H_ = max_H, *max_t_ = max_t;
}

int64_t orc_dp_cells(int qlen, int tlen, int w)
{
	int64_t n = 0;
	int r;
	if (w < 0) w = tlen > qlen ? tlen : qlen;
	for (r = 0; r < qlen + tlen - 1; ++r) {
		int st = 0, en = tlen - 1;
		if (st < r - qlen + 1) st = r - qlen + 1;
		if (en > r) en = r;
		if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
		if (en > (r + w) >> 1) en = (r + w) >> 1;
		if (st > en) break;
		n += en - st + 1;
	}
	return n;
}

/* ksw2_extd2_sse.c:27-403.  Returns the number of cells of the exact band processed. */
int orc_extd2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
              int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus, int flag,
              orc_ez_t *ez, uint32_t *cigar, int cigar_cap)
{
	int r, t, n_col_, tlen_, qlen_, last_st, last_en, last_max_H = 0, max_sc, min_sc, long_thres, long_diff, cells = 0;
	int with_cigar = !(flag & 0x01), approx_max = !!(flag & 0x08), right = !!(flag & 0x02), generic = !!(flag & 0x04);
	int32_t *H = 0, H0 = 0, last_H0_t = 0, qe0 = q + e; /* scalar q+e is taken before the swap (ksw2_extd2_sse.c:60) */
	int8_t qe, qe2, sc_mch, sc_mis, sc_N, m1;
	uint8_t *mem, *sf, *qr, *p = 0;
	int8_t *u, *v, *x, *y, *x2, *y2, *s;
	int *off = 0, *off_end = 0;
	cig_t cg = { 0, cigar_cap, 0, cigar };

	reset_ez(ez);
	if (m <= 1 || qlen <= 0 || tlen <= 0) return 0;
	if (q2 + e2 < q + e) { int8_t tq = q; q = q2; q2 = tq; tq = e; e = e2; e2 = tq; }
	qe = (int8_t)(q + e), qe2 = (int8_t)(q2 + e2);
	sc_mch = mat[0], sc_mis = mat[1];
	sc_N = mat[m * m - 1] == 0 ? (int8_t)-e2 : mat[m * m - 1];
	m1 = (int8_t)(m - 1);
	if (w < 0) w = tlen > qlen ? tlen : qlen;
	tlen_ = (tlen + 15) / 16;
	n_col_ = qlen < tlen ? qlen : tlen;
	n_col_ = ((n_col_ < w + 1 ? n_col_ : w + 1) + 15) / 16 + 1;
	qlen_ = (qlen + 15) / 16;
	for (t = 1, max_sc = mat[0], min_sc = mat[1]; t < m * m; ++t) {
		max_sc = max_sc > mat[t] ? max_sc : mat[t];
		min_sc = min_sc < mat[t] ? min_sc : mat[t];
	}
	if (-min_sc > 2 * (q + e)) return 0;
	long_thres = e != e2 ? (q2 - q) / (e - e2) - 1 : 0;
	if (q2 + e2 + long_thres * e2 > q + e + long_thres * e) ++long_thres;
	long_diff = long_thres * (e - e2) - (q2 - q) - e2;

	/* one block laid out like the reference: u v x y x2 y2 s sf (tlen_*16 each) + qr */
	mem = (uint8_t*)calloc((size_t)(tlen_ * 8 + qlen_ + 1) * 16, 1);
	u = (int8_t*)mem; v = u + tlen_ * 16; x = v + tlen_ * 16; y = x + tlen_ * 16;
	x2 = y + tlen_ * 16; y2 = x2 + tlen_ * 16; s = y2 + tlen_ * 16;
	sf = (uint8_t*)(s + tlen_ * 16); qr = sf + tlen_ * 16;
	memset(u, -q - e, (size_t)tlen_ * 16 * 4);
	memset(x2, -q2 - e2, (size_t)tlen_ * 16 * 2);
	if (!approx_max) {
		H = (int32_t*)malloc((size_t)tlen_ * 16 * 4);
		for (t = 0; t < tlen_ * 16; ++t) H[t] = NEG_INF;
	}
	if (with_cigar) {
		p = (uint8_t*)malloc((size_t)(qlen + tlen - 1) * n_col_ * 16 + 16);
		off = (int*)malloc((size_t)(qlen + tlen - 1) * sizeof(int) * 2);
		off_end = off + qlen + tlen - 1;
	}
	for (t = 0; t < qlen; ++t) qr[t] = query[qlen - 1 - t];
	memcpy(sf, target, tlen);

	for (r = 0, last_st = last_en = -1; r < qlen + tlen - 1; ++r) {
		int st = 0, en = tlen - 1, st0, en0, i;
		int8_t x1, x21, v1;
		const uint8_t *qrr = qr + (qlen - 1 - r);
		if (st < r - qlen + 1) st = r - qlen + 1;
		if (en > r) en = r;
		if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
		if (en > (r + w) >> 1) en = (r + w) >> 1;
		if (st > en) { ez->zdropped = 1; break; }
		st0 = st, en0 = en;
		cells += en0 - st0 + 1;
		st = st / 16 * 16, en = (en + 16) / 16 * 16 - 1;
		if (st > 0) {
			if (st - 1 >= last_st && st - 1 <= last_en) x1 = x[st - 1], x21 = x2[st - 1], v1 = v[st - 1];
			else x1 = -q - e, x21 = -q2 - e2, v1 = -q - e;
		} else {
			x1 = -q - e, x21 = -q2 - e2;
			v1 = r == 0 ? -q - e : r < long_thres ? -e : r == long_thres ? long_diff : -e2;
		}
		if (en >= r) {
			y[r] = -q - e, y2[r] = -q2 - e2;
			u[r] = r == 0 ? -q - e : r < long_thres ? -e : r == long_thres ? long_diff : -e2;
		}
		/* scores: 16-byte groups from st0, may run past en0 (and past s[] into sf[]) */
		if (!generic) {
			for (t = st0; t <= en0; t += 16) {
				int8_t tmp[16];
				for (i = 0; i < 16; ++i) {
					int8_t a = (int8_t)sf[t + i], b = (int8_t)qrr[t + i];
					tmp[i] = (a == m1 || b == m1) ? sc_N : a == b ? sc_mch : sc_mis;
				}
				memcpy(s + t, tmp, 16);
			}
		} else for (t = st0; t <= en0; ++t) s[t] = mat[sf[t] * m + qrr[t]];
		{ /* core: every cell of the aligned range [st,en] from the previous anti-diagonal */
			int n = en - st + 1;
			int8_t *ox = (int8_t*)malloc(n * 6), *ov = ox + n, *ox2 = ov + n, *ou = ox2 + n, *oy = ou + n, *oy2 = oy + n;
			uint8_t *pr = with_cigar ? p + (size_t)r * n_col_ * 16 : 0;
			if (with_cigar) off[r] = st, off_end[r] = en;
			for (i = 0; i < n; ++i) {
				int tt = st + i;
				ox[i] = i == 0 ? x1 : x[tt - 1];
				ov[i] = i == 0 ? v1 : v[tt - 1];
				ox2[i] = i == 0 ? x21 : x2[tt - 1];
				ou[i] = u[tt]; oy[i] = y[tt]; oy2[i] = y2[tt];
			}
			for (i = 0; i < n; ++i) {
				int tt = st + i;
				int8_t z = s[tt], vt1 = ov[i], ut = ou[i], tmp, d = 0;
				int8_t a = (int8_t)(ox[i] + vt1), b = (int8_t)(oy[i] + ut);
				int8_t a2 = (int8_t)(ox2[i] + vt1), b2 = (int8_t)(oy2[i] + ut);
				if (!with_cigar) {
					z = i8max(z, a); z = i8max(z, b); z = i8max(z, a2); z = i8max(z, b2);
				} else if (!right) {
					d = a > z ? 1 : 0;      z = i8max(z, a);
					d = b > z ? 2 : d;      z = i8max(z, b);
					d = a2 > z ? 3 : d;     z = i8max(z, a2);
					d = b2 > z ? 4 : d;     z = i8max(z, b2);
				} else {
					d = z > a ? 0 : 1;      z = i8max(z, a);
					d = z > b ? d : 2;      z = i8max(z, b);
					d = z > a2 ? d : 3;     z = i8max(z, a2);
					d = z > b2 ? d : 4;     z = i8max(z, b2);
				}
				z = i8min(z, sc_mch);
				u[tt] = (int8_t)(z - vt1);
				v[tt] = (int8_t)(z - ut);
				tmp = (int8_t)(z - q);  a = (int8_t)(a - tmp); b = (int8_t)(b - tmp);
				tmp = (int8_t)(z - q2); a2 = (int8_t)(a2 - tmp); b2 = (int8_t)(b2 - tmp);
				if (!with_cigar || !right) {
					x[tt]  = (int8_t)((a > 0 ? a : 0) - qe);
					y[tt]  = (int8_t)((b > 0 ? b : 0) - qe);
					x2[tt] = (int8_t)((a2 > 0 ? a2 : 0) - qe2);
					y2[tt] = (int8_t)((b2 > 0 ? b2 : 0) - qe2);
					if (with_cigar) d |= (a > 0 ? 0x08 : 0) | (b > 0 ? 0x10 : 0) | (a2 > 0 ? 0x20 : 0) | (b2 > 0 ? 0x40 : 0);
				} else {
					x[tt]  = (int8_t)((a >= 0 ? a : 0) - qe);
					y[tt]  = (int8_t)((b >= 0 ? b : 0) - qe);
					x2[tt] = (int8_t)((a2 >= 0 ? a2 : 0) - qe2);
					y2[tt] = (int8_t)((b2 >= 0 ? b2 : 0) - qe2);
					d |= (a >= 0 ? 0x08 : 0) | (b >= 0 ? 0x10 : 0) | (a2 >= 0 ? 0x20 : 0) | (b2 >= 0 ? 0x40 : 0);
				}
				if (with_cigar) pr[i] = (uint8_t)d;
			}
			free(ox);
		}
		if (!approx_max) {
			int32_t max_H, max_t;
			if (r > 0) exact_max(H, (uint8_t*)v, (uint8_t*)u, 1, st0, en0, 0, &max_H, &max_t);
			else H[0] = v[0] - qe0, max_H = H[0], max_t = 0;
			if (en0 == tlen - 1 && H[en0] > ez->mte) ez->mte = H[en0], ez->mte_q = r - en0;
			if (r - st0 == qlen - 1 && H[st0] > ez->mqe) ez->mqe = H[st0], ez->mqe_t = st0;
			if (apply_zdrop(ez, max_H, r, max_t, zdrop, e2)) break;
			if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H[tlen - 1];
			if (flag & 0x40) { /* ksw2_extd2_sse.c:361-367 */
				int32_t rH = last_max_H > max_H ? last_max_H : max_H;
				int32_t rq = qlen - (r - st0), rt = tlen - en0;
				int32_t rm = rq >= tlen - st0 ? tlen - st0 : rt >= qlen - (r - en0) ? qlen - (r - en0) : tlen + qlen - 1 - r;
				if (rH + rm * max_sc + end_bonus < ez->max) break;
			}
			last_max_H = max_H;
		} else {
			if (r > 0) {
				if (last_H0_t >= st0 && last_H0_t <= en0 && last_H0_t + 1 >= st0 && last_H0_t + 1 <= en0) {
					int32_t d0 = v[last_H0_t], d1 = u[last_H0_t + 1];
					if (d0 > d1) H0 += d0;
					else H0 += d1, ++last_H0_t;
				} else if (last_H0_t >= st0 && last_H0_t <= en0) {
					H0 += v[last_H0_t];
				} else {
					++last_H0_t, H0 += u[last_H0_t];
				}
			} else H0 = v[0] - qe0, last_H0_t = 0;
			if ((flag & 0x10) && apply_zdrop(ez, H0, r, last_H0_t, zdrop, e2)) break;
			if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H0;
		}
		last_st = st, last_en = en;
	}
	if (with_cigar) {
		int rev = !!(flag & 0x80);
		if (!ez->zdropped && !(flag & 0x40)) {
			backtrack(&cg, rev, p, off, off_end, (size_t)n_col_ * 16, tlen - 1, qlen - 1);
		} else if ((flag & 0x40) && ez->mqe + end_bonus > ez->max) {
			ez->reach_end = 1;
			backtrack(&cg, rev, p, off, off_end, (size_t)n_col_ * 16, ez->mqe_t, qlen - 1);
		} else if (ez->max_t >= 0 && ez->max_q >= 0) {
			backtrack(&cg, rev, p, off, off_end, (size_t)n_col_ * 16, ez->max_t, ez->max_q);
		}
		ez->n_cigar = cg.n;
	}
	free(mem); free(H); free(p); free(off);
	return cg.overflow ? -1 : cells;
}

/* ksw2_extz2_sse.c:16-296 (SSE4.1 code paths).  Returns the number of exact-band cells processed. */
int orc_extz2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
              int8_t q, int8_t e, int w, int zdrop, int end_bonus, int flag,
              orc_ez_t *ez, uint32_t *cigar, int cigar_cap)
{
	int r, t, n_col_, tlen_, qlen_, last_st, last_en, max_sc, min_sc, qe = q + e, cells = 0;
	int with_cigar = !(flag & 0x01), approx_max = !!(flag & 0x08), right = !!(flag & 0x02), generic = !!(flag & 0x04);
	int32_t *H = 0, H0 = 0, last_H0_t = 0;
	int8_t qe2_, sc_mch, sc_mis, sc_N, m1, max_sc_;
	uint8_t *mem, *sf, *qr, *p = 0, *u, *v, *x, *y, *s;
	int *off = 0, *off_end = 0;
	cig_t cg = { 0, cigar_cap, 0, cigar };

	reset_ez(ez);
	if (m <= 0 || qlen <= 0 || tlen <= 0) return 0;
	qe2_ = (int8_t)((q + e) * 2);
	sc_mch = mat[0], sc_mis = mat[1];
	sc_N = mat[m * m - 1] == 0 ? (int8_t)-e : mat[m * m - 1];
	m1 = (int8_t)(m - 1);
	max_sc_ = (int8_t)(mat[0] + (q + e) * 2);
	if (w < 0) w = tlen > qlen ? tlen : qlen;
	tlen_ = (tlen + 15) / 16;
	n_col_ = qlen < tlen ? qlen : tlen;
	n_col_ = ((n_col_ < w + 1 ? n_col_ : w + 1) + 15) / 16 + 1;
	qlen_ = (qlen + 15) / 16;
	for (t = 1, max_sc = mat[0], min_sc = mat[1]; t < m * m; ++t) {
		max_sc = max_sc > mat[t] ? max_sc : mat[t];
		min_sc = min_sc < mat[t] ? min_sc : mat[t];
	}
	if (-min_sc > 2 * (q + e)) return 0;

	mem = (uint8_t*)calloc((size_t)(tlen_ * 6 + qlen_ + 1) * 16, 1);
	u = mem; v = u + tlen_ * 16; x = v + tlen_ * 16; y = x + tlen_ * 16; s = y + tlen_ * 16;
	sf = s + tlen_ * 16; qr = sf + tlen_ * 16;
	if (!approx_max) {
		H = (int32_t*)malloc((size_t)tlen_ * 16 * 4);
		for (t = 0; t < tlen_ * 16; ++t) H[t] = NEG_INF;
	}
	if (with_cigar) {
		p = (uint8_t*)malloc((size_t)(qlen + tlen - 1) * n_col_ * 16 + 16);
		off = (int*)malloc((size_t)(qlen + tlen - 1) * sizeof(int) * 2);
		off_end = off + qlen + tlen - 1;
	}
	for (t = 0; t < qlen; ++t) qr[t] = query[qlen - 1 - t];
	memcpy(sf, target, tlen);

	for (r = 0, last_st = last_en = -1; r < qlen + tlen - 1; ++r) {
		int st = 0, en = tlen - 1, st0, en0, i;
		int8_t x1, v1;
		const uint8_t *qrr = qr + (qlen - 1 - r);
		if (st < r - qlen + 1) st = r - qlen + 1;
		if (en > r) en = r;
		if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
		if (en > (r + w) >> 1) en = (r + w) >> 1;
		if (st > en) { ez->zdropped = 1; break; }
		st0 = st, en0 = en;
		cells += en0 - st0 + 1;
		st = st / 16 * 16, en = (en + 16) / 16 * 16 - 1;
		if (st > 0) {
			if (st - 1 >= last_st && st - 1 <= last_en) x1 = (int8_t)x[st - 1], v1 = (int8_t)v[st - 1];
			else x1 = v1 = 0;
		} else x1 = 0, v1 = r ? q : 0;
		if (en >= r) y[r] = 0, u[r] = r ? (uint8_t)q : 0;
		if (!generic) {
			for (t = st0; t <= en0; t += 16) {
				int8_t tmp[16];
				for (i = 0; i < 16; ++i) {
					int8_t a = (int8_t)sf[t + i], b = (int8_t)qrr[t + i];
					tmp[i] = (a == m1 || b == m1) ? sc_N : a == b ? sc_mch : sc_mis;
				}
				memcpy(s + t, tmp, 16);
			}
		} else for (t = st0; t <= en0; ++t) s[t] = (uint8_t)mat[sf[t] * m + qrr[t]];
		{
			int n = en - st + 1;
			uint8_t *ox = (uint8_t*)malloc(n * 4), *ov = ox + n, *ou = ov + n, *oy = ou + n;
			uint8_t *pr = with_cigar ? p + (size_t)r * n_col_ * 16 : 0;
			if (with_cigar) off[r] = st, off_end[r] = en;
			for (i = 0; i < n; ++i) {
				int tt = st + i;
				/* x1/v1 enter through _mm_cvtsi32_si128 of an int8: a negative value also sets
				 * bytes 1..3 of the first 16-cell group (ksw2_extz2_sse.c:138-139) */
				if (i == 0) ox[i] = (uint8_t)x1, ov[i] = (uint8_t)v1;
				else {
					ox[i] = x[tt - 1], ov[i] = v[tt - 1];
					if (i < 4 && x1 < 0) ox[i] = 0xff;
					if (i < 4 && v1 < 0) ov[i] = 0xff;
				}
				ou[i] = u[tt], oy[i] = y[tt];
			}
			for (i = 0; i < n; ++i) {
				int tt = st + i;
				uint8_t vt1 = ov[i], ut = ou[i], d = 0;
				int8_t z = (int8_t)(s[tt] + (uint8_t)qe2_), a = (int8_t)(ox[i] + vt1), b = (int8_t)(oy[i] + ut), zz;
				uint8_t zu;
				if (!with_cigar) {
					z = i8max(z, a);
				} else if (!right) {
					d = a > z ? 1 : 0;
					z = i8max(z, a);
					d = b > z ? 2 : d;
				} else {
					d = z > a ? 0 : 1;
					z = i8max(z, a);
					d = z > b ? d : 2;
				}
				zu = u8max((uint8_t)z, (uint8_t)b);
				zu = u8min(zu, (uint8_t)max_sc_);
				u[tt] = (uint8_t)(zu - vt1);
				v[tt] = (uint8_t)(zu - ut);
				zz = (int8_t)(zu - (uint8_t)q);
				a = (int8_t)(a - zz); b = (int8_t)(b - zz);
				if (!with_cigar) {
					x[tt] = (uint8_t)i8max(a, 0); y[tt] = (uint8_t)i8max(b, 0);
				} else if (!right) {
					x[tt] = a > 0 ? (uint8_t)a : 0; d |= a > 0 ? 0x08 : 0;
					y[tt] = b > 0 ? (uint8_t)b : 0; d |= b > 0 ? 0x10 : 0;
				} else {
					x[tt] = 0 > a ? 0 : (uint8_t)a; d |= 0 > a ? 0 : 0x08;
					y[tt] = 0 > b ? 0 : (uint8_t)b; d |= 0 > b ? 0 : 0x10;
				}
				if (with_cigar) pr[i] = d;
			}
			free(ox);
		}
		if (!approx_max) {
			int32_t max_H, max_t;
			if (r > 0) exact_max(H, v, u, 0, st0, en0, qe, &max_H, &max_t);
			else H[0] = v[0] - qe - qe, max_H = H[0], max_t = 0;
			if (en0 == tlen - 1 && H[en0] > ez->mte) ez->mte = H[en0], ez->mte_q = r - en0;
			if (r - st0 == qlen - 1 && H[st0] > ez->mqe) ez->mqe = H[st0], ez->mqe_t = st0;
			if (apply_zdrop(ez, max_H, r, max_t, zdrop, e)) break;
			if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H[tlen - 1];
		} else {
			if (r > 0) {
				if (last_H0_t >= st0 && last_H0_t <= en0 && last_H0_t + 1 >= st0 && last_H0_t + 1 <= en0) {
					int32_t d0 = v[last_H0_t] - qe, d1 = u[last_H0_t + 1] - qe;
					if (d0 > d1) H0 += d0;
					else H0 += d1, ++last_H0_t;
				} else if (last_H0_t >= st0 && last_H0_t <= en0) {
					H0 += v[last_H0_t] - qe;
				} else {
					++last_H0_t, H0 += u[last_H0_t] - qe;
				}
				if ((flag & 0x10) && apply_zdrop(ez, H0, r, last_H0_t, zdrop, e)) break;
			} else H0 = v[0] - qe - qe, last_H0_t = 0;
			if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H0;
		}
		last_st = st, last_en = en;
	}
	if (with_cigar) {
		int rev = !!(flag & 0x80);
		if (!ez->zdropped && !(flag & 0x40)) {
			backtrack(&cg, rev, p, off, off_end, (size_t)n_col_ * 16, tlen - 1, qlen - 1);
		} else if ((flag & 0x40) && ez->mqe + end_bonus > ez->max) {
			ez->reach_end = 1;
			backtrack(&cg, rev, p, off, off_end, (size_t)n_col_ * 16, ez->mqe_t, qlen - 1);
		} else if (ez->max_t >= 0 && ez->max_q >= 0) {
			backtrack(&cg, rev, p, off, off_end, (size_t)n_col_ * 16, ez->max_t, ez->max_q);
		}
		ez->n_cigar = cg.n;
	}
	free(mem); free(H); free(p); free(off);
	return cg.overflow ? -1 : cells;
}

/* mb_align_pair (align.c:316-390): the generic-scoring switch (:334-335), the ungapped extension
 * fast path for EXTZ_ONLY calls with tlen >= qlen (:336-350), the ungapped global fast path for
 * qlen == tlen (:351-362), the band cap from the mismatch count when the shorter side is under 100
 * (:364-368, max_bw_from_mm :321-327), the max_sw_mat skip (:370-372) and the dispatch to extz2
 * (single affine) or extd2 (:373-378).  p = {a, b, b_ts, q, e, q2, e2, meth}.  Returns the DP's cells
 * (0 on a fast path or a skip), -1 on CIGAR overflow. */
int orc_align_pair(const int32_t p[8], int64_t max_sw_mat, int qlen, const uint8_t *qseq, int tlen, const uint8_t *tseq,
                   const int8_t *mat, int w, int end_bonus, int zdrop, int flag, orc_ez_t *ez, uint32_t *cigar, int cigar_cap)
{
	const int32_t a = p[0], b = p[1], b_ts = p[2], q = p[3], e = p[4], q2 = p[5], e2 = p[6], meth = p[7];
	int32_t j, n_mm = -1;
	if ((b_ts != 0 && b != b_ts) || meth) flag |= 0x04;
	if ((flag & 0x40) && tlen >= qlen) {
		reset_ez(ez);
		ez->score = ez->max = 0;
		for (j = 0; j < qlen; ++j) {
			const int32_t sc = mat[tseq[j] * 5 + qseq[j]];
			ez->score += sc;
			n_mm += tseq[j] > 3 || qseq[j] > 3 || sc < 0;
			if (ez->max < ez->score) ez->max = ez->score, ez->max_q = ez->max_t = j;
		}
		if (n_mm <= 2) {
			ez->mqe = ez->score, ez->mqe_t = qlen - 1;
			if (ez->mqe + end_bonus >= ez->max) {
				ez->reach_end = 1;
				if (cigar_cap < 1) return -1;
				cigar[0] = (uint32_t)qlen << 4, ez->n_cigar = 1;   /* ksw_push_cigar, a zero length included */
				return 0;
			}
		}
	} else if (qlen == tlen && !(flag & 0x40)) {
		const int32_t max_gapped = (qlen - 2) * a - 2 * (q + e);
		reset_ez(ez);
		for (j = 0, ez->score = 0; j < qlen; ++j) {
			const int32_t sc = mat[tseq[j] * 5 + qseq[j]];
			ez->score += sc;
			n_mm += tseq[j] > 3 || qseq[j] > 3 || sc < 0;
		}
		if (n_mm <= 3 || ez->score > max_gapped) {
			if (cigar_cap < 1) return -1;
			cigar[0] = (uint32_t)qlen << 4, ez->n_cigar = 1;
			return 0;
		}
	}
	if (n_mm >= 0 && (qlen < tlen ? qlen : tlen) < 100) {
		const int32_t x = n_mm * (a + b);
		int32_t max1 = 0, max2 = 0, mbw;
		if (x >= q + e) max1 = (x - q + e - 1) / e;
		if (x >= q2 + e2) max2 = (x - q2 + e2 - 1) / e2;
		mbw = max1 > max2 ? max1 : max2;
		if (w > mbw + 4) w = mbw + 4;
	}
	if (max_sw_mat > 0 && (int64_t)tlen * qlen > max_sw_mat) {
		reset_ez(ez);
		ez->zdropped = 1;
		return 0;
	}
	if (q == q2 && e == e2)
		return orc_extz2(qlen, qseq, tlen, tseq, 5, mat, (int8_t)q, (int8_t)e, w, zdrop * a, end_bonus, flag, ez, cigar, cigar_cap);
	return orc_extd2(qlen, qseq, tlen, tseq, 5, mat, (int8_t)q, (int8_t)e, (int8_t)q2, (int8_t)e2, w, zdrop * a, end_bonus, flag, ez,
	                 cigar, cigar_cap);
}

/* ksw_ll_qinit + ksw_ll_u8_core / ksw_ll_i16_core (ksw2_ll_sse.c:49-333): local Smith-Waterman
 * with affine gaps, score and end positions only, as the mate-rescue step calls it (pe.c:260-272).
 *
 * Restated over the query in plain left-to-right order instead of the SSE striped layout; the
 * observable effects of the striping are kept:
 *   - the query is padded to p*slen positions (p = 16 bytes or 8 words per vector,
 *     slen = ceil(qlen/p)); padded positions score -1 (ksw2_ll_sse.c:82,90) and take part in the
 *     row maxima;
 *   - the first pass carries F only inside each stripe of slen consecutive positions (the
 *     vector lanes, :149-172); E for the next row is taken from that first-pass H (:162-166);
 *   - the lazy-F loop (:174-185) makes H the fixed point of the full-column F chain, sourced from
 *     the first-pass H; here that chain is one sequential pass over the row;
 *   - the row maximum is taken over the first-pass H (:159,187); qe is the smallest position
 *     holding the maximum of the post-lazy-F row kept at te (:200-204);
 *   - u8 arithmetic saturates (adds/subs with the matrix biased by shift, :154-170) and the
 *     scan stops once gmax+shift reaches 255 (:197), reporting 255 without qe or a second best;
 *   - the second best walks the b array of per-row maxima (:188-193), merged only into an entry
 *     whose stored row is the previous row, outside te +- ceil(score / max(mat)) (:205-213).
 * Returns 0, or -1 for arguments not restated: m outside 1..16, or gapo < 1.  With gapo == 0 the
 * lazy-F loop's early exit (f <= H - (gapo+gape) after f -= gape) also fires on a tie, before
 * the fixed point, at a point set by the SSE lane order; for gapo >= 1 it fires only once no
 * lane can change, so the fixed point above is what the reference computes. */
static inline int ll_sat_u8(int x) { return x < 0 ? 0 : x > 255 ? 255 : x; }
static inline int ll_sat_i16(int x) { return x < -32768 ? -32768 : x > 32767 ? 32767 : x; }

int orc_ll(int size, int qlen, const uint8_t *query, int m, const int8_t *mat, int tlen, const uint8_t *target,
           int gapo, int gape, int xtra, orc_llrst_t *r)
{
	int p, slen, n, i, k, te = -1, gmax = 0, minsc, endsc, u8, shift, qmax, gapoe_, gape_, nb = 0, mb = 0;
	int *Hp, *Hc, *Hd, *E, *Hmax, *sc;
	uint64_t *b = 0;
	uint8_t sh, md;

	r->score = 0, r->te = r->qe = r->score2 = r->te2 = -1;
	if (m < 1 || m > 16 || gapo < 1) return -1;
	u8 = size <= 1;
	p = u8 ? 16 : 8;
	slen = (qlen + p - 1) / p;
	n = slen * p;
	for (k = 0, sh = 127, md = 0; k < m * m; ++k) {     /* :67-74 in uint8 arithmetic */
		if (mat[k] < (int8_t)sh) sh = (uint8_t)mat[k];
		if (mat[k] > (int8_t)md) md = (uint8_t)mat[k];
	}
	qmax = md;
	shift = (uint8_t)(256 - sh);
	minsc = (xtra & 0x40000) ? xtra & 0xffff : 0x10000;
	endsc = (xtra & 0x20000) ? xtra & 0xffff : 0x10000;
	/* the SSE code broadcasts gapo+gape and gape into bytes or words */
	gapoe_ = u8 ? (uint8_t)(gapo + gape) : (uint16_t)(gapo + gape);
	gape_ = u8 ? (uint8_t)gape : (uint16_t)gape;

	sc = (int*)malloc(sizeof(int) * ((size_t)m * n + 1));
	Hp = (int*)calloc((size_t)n + 1, sizeof(int));      /* previous row, post-lazy-F */
	Hc = (int*)calloc((size_t)n + 1, sizeof(int));      /* current row, post-lazy-F */
	Hd = (int*)calloc((size_t)n + 1, sizeof(int));      /* current row, first pass */
	E = (int*)calloc((size_t)n + 1, sizeof(int));
	Hmax = (int*)calloc((size_t)n + 1, sizeof(int));
	for (i = 0; i < m; ++i)                             /* query profile, :77-94 */
		for (k = 0; k < n; ++k) {
			int s = k >= qlen ? -1 : mat[i * m + query[k]];
			sc[i * n + k] = u8 ? (uint8_t)(int8_t)(s + shift) : s;
		}
	for (i = 0; i < tlen; ++i) {
		int f = 0, imax = 0, *t;
		const int *S = sc + (size_t)target[i] * n;
		for (k = 0; k < n; ++k) {
			int h, e = E[k], d = k == 0 ? 0 : Hp[k - 1];
			if (k % slen == 0) f = 0;                   /* a new vector lane starts its own F */
			if (u8) h = ll_sat_u8(ll_sat_u8(d + S[k]) - shift);
			else h = ll_sat_i16(d + S[k]);
			h = h > e ? h : e;
			h = h > f ? h : f;
			Hd[k] = h;
			imax = imax > h ? imax : h;
			if (u8) {
				e = ll_sat_u8(e - gape_), f = ll_sat_u8(f - gape_);
				E[k] = e > ll_sat_u8(h - gapoe_) ? e : ll_sat_u8(h - gapoe_);
				f = f > ll_sat_u8(h - gapoe_) ? f : ll_sat_u8(h - gapoe_);
			} else {                                    /* _mm_subs_epu16 on words, then signed max */
				int a1, a2;
				a1 =(int16_t)(uint16_t)((uint16_t)e > (uint16_t)gape_ ? (uint16_t)e - (uint16_t)gape_ : 0);
				a2 = (int16_t)(uint16_t)((uint16_t)h > (uint16_t)gapoe_ ? (uint16_t)h - (uint16_t)gapoe_ : 0);
				E[k] = a1 > a2 ? a1 : a2;
				a1 = (int16_t)(uint16_t)((uint16_t)f > (uint16_t)gape_ ? (uint16_t)f - (uint16_t)gape_ : 0);
				f = a1 > a2 ? a1 : a2;
			}
		}
		for (k = 0, f = 0; k < n; ++k) {               /* the lazy-F fixed point */
			int h = Hd[k] > f ? Hd[k] : f, a;
			Hc[k] = h;
			if (u8) {
				a = ll_sat_u8(Hd[k] - gapoe_);
				f = ll_sat_u8(f - gape_);
			} else {
				a = (int16_t)(uint16_t)((uint16_t)Hd[k] > (uint16_t)gapoe_ ? (uint16_t)Hd[k] - (uint16_t)gapoe_ : 0);
				f = (int16_t)(uint16_t)((uint16_t)f > (uint16_t)gape_ ? (uint16_t)f - (uint16_t)gape_ : 0);
			}
			f = f > a ? f : a;
		}
		if (imax >= minsc) {                            /* :188-193 */
			if (nb == 0 || (int32_t)b[nb - 1] + 1 != i) {
				if (nb == mb) {
					mb = mb ? mb * 2 : 16;
					b = (uint64_t*)realloc(b, sizeof(uint64_t) * mb);
				}
				b[nb++] = (uint64_t)imax << 32 | (uint32_t)i;
			} else if ((int)(b[nb - 1] >> 32) < imax) b[nb - 1] = (uint64_t)imax << 32 | (uint32_t)i;
		}
		if (imax > gmax) {
			gmax = imax, te = i;
			memcpy(Hmax, Hc, sizeof(int) * n);
			if ((u8 && gmax + shift >= 255) || gmax >= endsc) break;
		}
		t = Hp, Hp = Hc, Hc = t;
	}
	r->score = u8 && gmax + shift >= 255 ? 255 : gmax;
	r->te = te;
	if (!u8 || r->score != 255) {
		int best = -1, lo, hi;
		for (k = 0; k < n; ++k)                         /* smallest position holding the maximum */
			if (Hmax[k] > best) best = Hmax[k], r->qe = k;
		if (b) {
			int L = (r->score + qmax - 1) / qmax;
			lo = te - L, hi = te + L;
			for (i = 0; i < nb; ++i) {
				int e = (int32_t)b[i];
				if ((e < lo || e > hi) && (int)(b[i] >> 32) > r->score2)
					r->score2 = (int)(b[i] >> 32), r->te2 = e;
			}
		}
	}
	free(b); free(sc); free(Hp); free(Hc); free(Hd); free(E); free(Hmax);
	return 0;
}

/* ksw_gen_nt4_mat (ksw2.h:199-219): 5x5 scores for A,C,G,T,N; mt 1 = C->T, 2 = G->A conversion */
static void gen_nt4_mat(int8_t *mat, int a, int b, int b_ts, int b_ambi, int mt)
{
	int i, j;
	a = a < 0 ? -a : a, b = b > 0 ? -b : b, b_ambi = b_ambi > 0 ? -b_ambi : b_ambi;
	for (i = 0; i < 4; ++i) {
		for (j = 0; j < 4; ++j) mat[i * 5 + j] = (int8_t)(i == j ? a : b);
		mat[i * 5 + 4] = (int8_t)b_ambi;
	}
	for (j = 0; j < 5; ++j) mat[20 + j] = (int8_t)b_ambi;
	if (b_ts != 0 && b_ts != b) {
		b_ts = b_ts > 0 ? -b_ts : b_ts;
		mat[2] = mat[8] = mat[10] = mat[16] = (int8_t)b_ts;
	}
	if (mt == 1) mat[8] = (int8_t)a;
	else if (mt == 2) mat[10] = (int8_t)a;
}

/* mb_ungap (pe.c:209-247): the offset (target position - query position) of most k-mer hits, each
 * target k-mer voting for the diagonal of the LAST query position where the same k-mer ends.
 * Bases go through the methylation conversion first (mt 1: C->T, 2: G->A); an N restarts the
 * k-mer.  Returns the top vote count; *max_i its first diagonal (-1 when no vote), *n_good the
 * diagonals with more than half of it, *n_kmer the distinct query k-mers.  kmer <= 8. */
int32_t orc_ungap(int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq, int32_t kmer, int32_t mt,
                  int32_t *max_i, int32_t *n_good, int32_t *n_kmer)
{
	static const uint8_t conv[3][4] = { {0, 1, 2, 3}, {0, 3, 2, 3}, {0, 1, 0, 3} };
	const uint8_t *cv = conv[mt == 1 ? 1 : mt == 2 ? 2 : 0];
	uint32_t mask = (1U << 2 * kmer) - 1, x = 0;
	int32_t i, l, best = 0, *a;
	uint16_t *h;
	*max_i = -1, *n_good = *n_kmer = 0;
	if (qlen >= 65535) return 0;
	a = (int32_t*)calloc((size_t)tlen + 1, sizeof(int32_t));
	h = (uint16_t*)calloc((size_t)mask + 1, sizeof(uint16_t));
	for (i = l = 0; i < qlen; ++i) {
		if (qseq[i] > 3) { x = 0, l = 0; continue; }
		x = (x << 2 | cv[qseq[i]]) & mask;
		if (++l < kmer) continue;
		*n_kmer += h[x] == 0;
		h[x] = (uint16_t)i;                         /* the last occurrence wins */
	}
	for (i = l = 0, x = 0; i < tlen; ++i) {
		if (tseq[i] > 3) { x = 0, l = 0; continue; }
		x = (x << 2 | cv[tseq[i]]) & mask;
		if (++l >= kmer && h[x] > 0 && i >= h[x]) ++a[i - h[x]];
	}
	for (i = 0; i < tlen; ++i)
		if (a[i] > best) best = a[i], *max_i = i;
	for (i = 0; i < tlen; ++i) *n_good += a[i] > best >> 1;
	free(a); free(h);
	return best;
}

/* mb_matesw_align (pe.c:249-312) up to the hit it fills, before mb_append_cigar/mb_update_extra:
 * ksw_ll with the match-1 matrix and gaps divided by a (u8 core when the best possible score stays
 * below 255 - mismatch), then ksw_extz2 of the reversed prefixes ending at (qe, te) with
 * EXTZ_ONLY|RIGHT|REV_CIGAR (and GENERIC_SC under a methylation conversion).
 * p = {a, b, b_ts, b_ambi, q, e, bw, zdrop, end_bonus, min_dp_max, min_len}.  Returns 0. */
int orc_matesw_align(const int32_t p[11], int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq, int32_t min_sc,
                     int32_t mt, orc_rescue_t *r, uint32_t *cigar, int cigar_cap)
{
	const int a = p[0], b = p[1], b_ts = p[2], b_ambi = p[3], q = p[4], e = p[5], bw = p[6], zdrop = p[7], eb = p[8];
	const int min_dp_max = p[9], min_len = p[10];
	int32_t max_sc = qlen < tlen ? qlen : tlen, b_mm = (b + a - 1) / a, sz, i;
	int8_t mat[25];
	memset(r, 0, sizeof(*r));
	r->ll.score = r->ll.te = r->ll.qe = r->ll.score2 = r->ll.te2 = -1;
	if (max_sc >= 32767) return 0;
	gen_nt4_mat(mat, 1, b_mm, (b_ts + a - 1) / a, (b_ambi + a - 1) / a, mt);
	sz = max_sc < 255 - b_mm ? 1 : 2;
	orc_ll(sz, qlen, qseq, 5, mat, tlen, tseq, (q + a - 1) / a, (e + a - 1) / a, 0x40000 | min_len, &r->ll);
	r->ran_ll = 1;
	if (r->ll.score >= min_dp_max && r->ll.score >= min_sc) {
		int32_t qe = r->ll.qe + 1, te = r->ll.te + 1, flag = 0x40 | 0x02 | 0x80 | (mt ? 0x04 : 0);
		uint8_t *rq = (uint8_t*)malloc((size_t)qe + 1), *rt = (uint8_t*)malloc((size_t)te + 1);
		for (i = 0; i < qe; ++i) rq[i] = qseq[qe - 1 - i];
		for (i = 0; i < te; ++i) rt[i] = tseq[te - 1 - i];
		gen_nt4_mat(mat, a, b, b_ts, b_ambi, mt);
		orc_extz2(qe, rq, te, rt, 5, mat, (int8_t)q, (int8_t)e, bw, zdrop, eb, flag, &r->ez, cigar, cigar_cap);
		r->ran_dp = 1;
		free(rq); free(rt);
		if (r->ez.n_cigar > 0 && r->ez.max >= min_dp_max * a) {
			r->rescued = 1, r->qe = qe, r->te = te;
			r->ts = te - (r->ez.reach_end ? r->ez.mqe_t + 1 : r->ez.max_t + 1);
			r->qs = qe - (r->ez.reach_end ? qe : r->ez.max_q + 1);
			r->dp_max2 = (int32_t)((double)r->ez.max / r->ll.score * r->ll.score2 + .499);
			if (r->dp_max2 < 0) r->dp_max2 = 0;
		}
	}
	return 0;
}

/* one rescue candidate of mb_matesw_core (pe.c:342-349), window-relative: mb_ungap over the whole
 * window, the narrowing to [max_i - qlen/2, +2*qlen) when the vote is clear, and mb_matesw_align on
 * the (narrowed) window when the vote passes.  ug = {max_ug, max_i, n_good, n_kmer}, win2 = {ts2, te2}. */
int orc_matesw_cand(const int32_t p[11], int32_t qlen, const uint8_t *qseq, int32_t wlen, const uint8_t *win, int32_t min_sc,
                    int32_t mt, int32_t ug[4], int32_t win2[2], int32_t *aligned, orc_rescue_t *r, uint32_t *cigar, int cigar_cap)
{
	int32_t ts2 = 0, te2 = wlen;
	ug[0] = orc_ungap(qlen, qseq, wlen, win, 7, mt, &ug[1], &ug[2], &ug[3]);
	if (ug[0] >= 10 && ug[0] >= qlen >> 1 && ug[2] == 1) {
		ts2 = ug[1] - qlen / 2, te2 = ts2 + qlen * 2;
		if (ts2 < 0) ts2 = 0;
		if (te2 > wlen) te2 = wlen;
	}
	win2[0] = ts2, win2[1] = te2;
	*aligned = ug[0] >= 10 || ug[0] >= ug[3] * 0.33;
	memset(r, 0, sizeof(*r));
	r->ll.score = r->ll.te = r->ll.qe = r->ll.score2 = r->ll.te2 = -1;
	if (*aligned) orc_matesw_align(p, qlen, qseq, te2 - ts2, win + ts2, min_sc, mt, r, cigar, cigar_cap);
	return 0;
}

/* mb_log2 (mbpriv.h:105-113): the float bit-pattern approximation of log2 */
static float orc_log2(float x)
{
	union { float f; uint32_t i; } z = { x };
	float l = (float)(((z.i >> 23) & 255u) - 128u);
	z.i &= ~(255u << 23);
	z.i += 127u << 23;
	l += (-0.34484843f * z.f + 2.02466578f) * z.f - 0.67487759f;
	return l;
}

/* mb_fix_cigar (align.c:70-142) and mb_update_extra (align.c:219-258) up to the scores it sets,
 * without the =/X, cs and MD strings.  qseq and tseq point at the hit's first bases (qs, ts), as
 * the reference's callers pass them.  The CIGAR is edited in place:
 *   - an indel between two M runs moves left while the bases before it equal its last ones;
 *   - a run of adjacent I and D operations becomes one I and one D;
 *   - zero-length operations are squeezed out and equal neighbours merged;
 *   - a leading I or D is dropped, moving qs (qe for a reverse hit) or ts.
 * Then blen, mlen and the ambiguous-base count, and dp_max, the best score of any prefix of the
 * alignment (never below 0), with gaps costing q+e, or q+e*log2(1+len) when log_gap. */
void orc_update_extra(orc_extra_t *r, uint32_t *cigar, const uint8_t *qseq, const uint8_t *tseq, const int8_t *mat, int8_t q, int8_t e,
                      int log_gap)
{
	int32_t toff = 0, qoff = 0, to_shrink = 0, qshift = 0, tshift = 0;
	uint32_t k, n = (uint32_t)r->n_cigar;
	double sc = 0.0, best = 0.0;
	if (n == 0) return;
	if (n > 1) {
		for (k = 0; k < n; ++k) {
			uint32_t op = cigar[k] & 0xf, len = cigar[k] >> 4;
			if (len == 0) to_shrink = 1;
			if (op == 0) {
				toff += len, qoff += len;
			} else if (op == 1 || op == 2) {
				if (k > 0 && k < n - 1 && (cigar[k - 1] & 0xf) == 0 && (cigar[k + 1] & 0xf) == 0) {
					int l, prev = (int)(cigar[k - 1] >> 4);
					const uint8_t *sq = op == 1 ? qseq : tseq;
					int off = op == 1 ? qoff : toff;
					for (l = 0; l < prev; ++l)
						if (sq[off - 1 - l] != sq[off + (int)len - 1 - l]) break;
					if (l > 0) cigar[k - 1] -= (uint32_t)l << 4, cigar[k + 1] += (uint32_t)l << 4, qoff -= l, toff -= l;
					if (l == prev) to_shrink = 1;
				}
				if (op == 1) qoff += len;
				else toff += len;
			} else if (op == 3) {
				toff += len;
			}
		}
		for (k = 0; k + 2 < n; ++k) {                  /* a run like 5I6D7I becomes 12I6D */
			if ((cigar[k] & 0xf) > 0 && (cigar[k] & 0xf) + (cigar[k + 1] & 0xf) == 3) {
				uint32_t l, cnt[16] = {0};
				for (l = k; l < n; ++l) {
					uint32_t op = cigar[l] & 0xf;
					if (op == 1 || op == 2 || cigar[l] >> 4 == 0) cnt[op] += cigar[l] >> 4;
					else break;
				}
				if (cnt[1] > 0 && cnt[2] > 0 && l - k > 2) {
					cigar[k] = cnt[1] << 4 | 1;
					cigar[k + 1] = cnt[2] << 4 | 2;
					for (k += 2; k < l; ++k) cigar[k] &= 0xf;
					to_shrink = 1;
				}
				k = l;
			}
		}
		if (to_shrink) {
			uint32_t l = 0;
			for (k = 0; k < n; ++k)
				if (cigar[k] >> 4 != 0) cigar[l++] = cigar[k];
			n = l;
			for (k = l = 0; k < n; ++k)
				if (k == n - 1 || (cigar[k] & 0xf) != (cigar[k + 1] & 0xf)) cigar[l++] = cigar[k];
				else cigar[k + 1] += cigar[k] >> 4 << 4;
			n = l;
		}
		if ((cigar[0] & 0xf) == 1 || (cigar[0] & 0xf) == 2) {
			int32_t l = (int32_t)(cigar[0] >> 4);
			if ((cigar[0] & 0xf) == 1) {
				if (r->rev) r->qe -= l;
				else r->qs += l;
				qshift = l;
			} else r->ts += l, tshift = l;
			--n;
			memmove(cigar, cigar + 1, n * 4);
		}
	}
	qseq += qshift, tseq += tshift;
	r->n_cigar = (int32_t)n;
	r->blen = r->mlen = r->n_ambi = 0;
	toff = qoff = 0;
	for (k = 0; k < n; ++k) {
		uint32_t op = cigar[k] & 0xf, len = cigar[k] >> 4, l;
		int na = 0, nd = 0;
		if (op == 0) {
			for (l = 0; l < len; ++l) {
				int cq = qseq[qoff + l], ct = tseq[toff + l];
				if (ct > 3 || cq > 3) ++na;
				else if (ct != cq) nd += mat[ct * 5 + cq] < 0;
				sc += mat[ct * 5 + cq];
				if (sc < 0) sc = 0;
				else best = best > sc ? best : sc;
			}
			r->blen += (int32_t)len - na, r->mlen += (int32_t)len - (na + nd), r->n_ambi += na;
			toff += len, qoff += len;
		} else if (op == 1 || op == 2) {
			const uint8_t *sq = op == 1 ? qseq + qoff : tseq + toff;
			for (l = 0; l < len; ++l) na += sq[l] > 3;
			r->blen += (int32_t)len - na, r->n_ambi += na;
			if (log_gap) sc -= q + (double)e * orc_log2((float)(1.0 + len));
			else sc -= q + e;
			if (sc < 0) sc = 0;
			if (op == 1) qoff += len;
			else toff += len;
		}
	}
	r->dp_max = (int32_t)(best + .499);
}

/* klib's radix sort (ksort.h:108-161 as seed.c:9-18 instantiates it): insertion sort up to 64
 * elements, else an in-place most-significant-byte-first radix sort, 8 bits a pass, whose cycle
 * permutation is not stable; buckets of 2-64 elements finish by insertion sort.  Equal keys end up
 * in the order this procedure leaves them, which anchor ties depend on.  key(i) reads element i. */
typedef struct { int64_t b, e; } rs_bucket_t;

static void rs_insert(void *a, size_t sz, int64_t beg, int64_t end, uint64_t (*key)(const void*))
{
	char tmp[64], *base = (char*)a;
	int64_t i, j;
	for (i = beg + 1; i < end; ++i)
		if (key(base + i * sz) < key(base + (i - 1) * sz)) {
			memcpy(tmp, base + i * sz, sz);
			for (j = i; j > beg && key(tmp) < key(base + (j - 1) * sz); --j) memcpy(base + j * sz, base + (j - 1) * sz, sz);
			memcpy(base + j * sz, tmp, sz);
		}
}

static void rs_msd(void *a, size_t sz, int64_t beg, int64_t end, int s, uint64_t (*key)(const void*))
{
	char tmp[64], swp[64], *base = (char*)a;
	rs_bucket_t b[256];
	int64_t i, k, l;
	for (k = 0; k < 256; ++k) b[k].b = b[k].e = beg;
	for (i = beg; i < end; ++i) ++b[key(base + i * sz) >> s & 255].e;
	for (k = 1; k < 256; ++k) b[k].e += b[k - 1].e - beg, b[k].b = b[k - 1].e;
	for (k = 0; k < 256;) {                               /* the in-place cycle permutation */
		if (b[k].b != b[k].e) {
			if ((l = (int64_t)(key(base + b[k].b * sz) >> s & 255)) != k) {
				memcpy(tmp, base + b[k].b * sz, sz);
				do {
					memcpy(swp, tmp, sz);
					memcpy(tmp, base + b[l].b * sz, sz);
					memcpy(base + b[l].b++ * sz, swp, sz);
					l = (int64_t)(key(tmp) >> s & 255);
				} while (l != k);
				memcpy(base + b[k].b++ * sz, tmp, sz);
			} else ++b[k].b;
		} else ++k;
	}
	for (b[0].b = beg, k = 1; k < 256; ++k) b[k].b = b[k - 1].e;
	if (s) {
		s = s > 8 ? s - 8 : 0;
		for (k = 0; k < 256; ++k)
			if (b[k].e - b[k].b > 64) rs_msd(a, sz, b[k].b, b[k].e, s, key);
			else if (b[k].e - b[k].b > 1) rs_insert(a, sz, b[k].b, b[k].e, key);
	}
}

static void rs_sort(void *a, size_t sz, int64_t n, uint64_t (*key)(const void*))
{
	if (n <= 64) rs_insert(a, sz, 0, n, key);
	else rs_msd(a, sz, 0, n, 56, key);
}

static uint64_t key_x0(const void *p) { return ((const orc_sai_t*)p)->x[0]; }
static uint64_t key_size(const void *p) { return ((const orc_sai_t*)p)->size; }
static uint64_t key_info(const void *p) { return ((const orc_sai_t*)p)->info; }
static uint64_t key_tpos(const void *p) { return (uint64_t)((const orc_anchor_t*)p)->tpos; }

/* mb_seed_sort_dedup (seed.c:109-138): by x[0]; a group of equal x[0] by size, reversed, and runs of
 * equal size by info; then one of each (x[0], size, info).  Returns the new count. */
static int64_t seed_sort_dedup(orc_sai_t *u, int64_t n)
{
	int64_t i, i0, j, k, k0;
	if (n <= 1) return n;
	rs_sort(u, sizeof(*u), n, key_x0);
	for (i = 1, i0 = 0; i <= n; ++i) {
		if (i == n || u[i].x[0] != u[i0].x[0]) {
			if (i - i0 > 1) {
				rs_sort(u + i0, sizeof(*u), i - i0, key_size);
				for (k = 0; k < (i - i0) / 2; ++k) {
					orc_sai_t t = u[i0 + k]; u[i0 + k] = u[i - 1 - k]; u[i - 1 - k] = t;
				}
				for (k = i0 + 1, k0 = i0; k <= i; ++k)
					if (k == i || u[k0].size != u[k].size) {
						if (k - k0 > 1) rs_sort(u + k0, sizeof(*u), k - k0, key_info);
						k0 = k;
					}
			}
			i0 = i;
		}
	}
	for (i = 1, j = 0; i < n; ++i)
		if (!(u[i].x[0] == u[j].x[0] && u[i].size == u[j].size && u[i].info == u[j].info)) u[++j] = u[i];
	return j + 1;
}

/* l2b_pos2cid + l2b_intv2cid (l2bit.c:9-34): the contig of [st,en) on the forward+reverse text */
static int64_t intv2cid(int64_t n_ctg, const uint64_t *ctg_off, const uint64_t *ctg_len, uint64_t tot_len, uint64_t st, uint64_t en,
                        int64_t *cst, int *rev)
{
	int64_t s, len = (int64_t)(en - st), lo = 0, hi = n_ctg, mid;
	if (en > tot_len * 2) return -3;
	if (st < tot_len && tot_len < en) return -2;
	*rev = st >= tot_len;
	s = st < tot_len ? (int64_t)st : (int64_t)(tot_len * 2 - en);
	while (lo < hi) {
		mid = (lo + hi) / 2;
		if ((int64_t)ctg_off[mid] <= s && s < (int64_t)(ctg_off[mid] + ctg_len[mid])) {
			*cst = s - (int64_t)ctg_off[mid];
			return s + len <= (int64_t)(ctg_off[mid] + ctg_len[mid]) ? mid : -1;
		} else if (s < (int64_t)ctg_off[mid]) hi = mid;
		else lo = mid + 1;
	}
	return -1;
}

/* mb_anchor (seed.c:268-339) without a methylation conversion: the read's intervals sorted and
 * deduplicated, SA positions of every interval of size <= max_occ or of max_occ evenly stepped rows
 * of a larger one (seed.c:304-315), each position of each interval sharing (x[0], size) mapped to its
 * contig and strand (seed.c:176-210), the anchors radix-sorted by concatenated tpos, tpos made
 * contig-local, and anchors contained in another with the same diagonal dropped (seed.c:140-166).
 * u is edited in place.  Returns the number of anchors written to v (cap entries at most; -1 when
 * more would be needed). */
int64_t orc_anchor(const orc_bwt_t *b, int64_t n_ctg, const uint64_t *ctg_off, const uint64_t *ctg_len, orc_sai_t *u, int64_t n_u,
                   int32_t qlen, int32_t max_occ, orc_anchor_t *v, int64_t cap)
{
	int64_t i, i0, j, k, n = 0, ng = 0;
	const uint64_t tot = ctg_off[n_ctg - 1] + ctg_len[n_ctg - 1];
	n_u = seed_sort_dedup(u, n_u);
	for (i = 1, i0 = 0; i <= n_u; ++i) {
		if (i < n_u && u[i].x[0] == u[i0].x[0] && u[i].size == u[i0].size) continue;
		/* one group: intervals [i0, i) share the rows to locate */
		{
			const orc_sai_t *q = &u[i0];
			int64_t jj, cnt = 0;
			for (jj = 0; jj < (int64_t)q->size && cnt < max_occ; ++cnt) {
				const uint64_t row = q->x[0] + jj, pos = orc_sa(b, row);
				if ((int64_t)q->size <= max_occ) ++jj;
				else {
					int32_t step = (int32_t)((q->size - (uint64_t)jj) / (uint64_t)(max_occ - cnt));   /* int32 in seed.c:310 */
					jj += step < 1 ? 1 : step;
				}
				for (j = i0; j < i; ++j) {
					const int32_t qs = (int32_t)(u[j].info >> 32), qe = (int32_t)u[j].info, len = qe - qs;
					int64_t cst = 0, tid;
					int rev = 0;
					tid = intv2cid(n_ctg, ctg_off, ctg_len, tot, pos, pos + len, &cst, &rev);
					if (tid < 0) continue;
					if (n >= cap) return -1;
					memset(&v[n], 0, sizeof(v[n]));
					v[n].sid = (int32_t)(tid << 1 | rev);
					v[n].len = len;
					v[n].qpos = rev ? qlen - 1 - qs : qs + len - 1;
					v[n].tpos = (int64_t)(ctg_off[tid] * 2 + ctg_len[tid] * rev) + cst + len - 1;
					++n;
				}
			}
		}
		++ng;
		i0 = i;
	}
	rs_sort(v, sizeof(*v), n, key_tpos);
	for (i = 0; i < n; ++i) v[i].tpos -= (int64_t)(ctg_off[v[i].sid >> 1] * 2 + ctg_len[v[i].sid >> 1] * (v[i].sid & 1));
	for (i = 1; i < n; ++i) {                              /* mb_anchor_dedup, seed.c:140-166 */
		orc_anchor_t *ai = &v[i];
		const int64_t tsi = ai->tpos + 1 - ai->len;
		const int32_t qsi = ai->qpos + 1 - ai->len;
		for (j = i - 1, k = 0; j >= 0 && k < 100; --j, ++k) {
			orc_anchor_t *aj = &v[j];
			int64_t tsj;
			int32_t qsj;
			if (aj->sid != ai->sid || aj->tpos < tsi) break;
			tsj = aj->tpos + 1 - aj->len, qsj = aj->qpos + 1 - aj->len;
			if (tsj >= tsi) {
				if (tsj - tsi == qsj - qsi) aj->flt = 1;
			} else if (ai->tpos == aj->tpos) {
				if (ai->qpos == aj->qpos) ai->flt = 1;
			}
		}
	}
	for (i = j = 0; i < n; ++i)
		if (!v[i].flt) v[j++] = v[i];
	(void)ng;
	return j;
}

/* ---- chaining: mb_lchain_dp (lchain.c:145-240) ---------------------------------------------- */

typedef struct { uint64_t x, y; } orc_u128_t;   /* mb128_t (mbpriv.h:42) */
static uint64_t key_u128x(const void *p) { return ((const orc_u128_t*)p)->x; }

/* comput_sc (lchain.c:125-143): the score of chaining aj -> ai, or INT32_MIN when they cannot chain */
static int32_t chain_sc(const orc_anchor_t *ai, const orc_anchor_t *aj, int32_t max_dist_x, int32_t max_dist_y, int32_t bw,
                        float pen_gap)
{
	const int64_t dq = (int64_t)ai->qpos - aj->qpos;
	int64_t dr, dd, dg, sc;
	if (dq <= 0 || dq > max_dist_y + ai->len) return INT32_MIN;
	if (ai->sid != aj->sid) return INT32_MIN;
	dr = ai->tpos - aj->tpos;
	if (dr <= 0 || dq > max_dist_x + ai->len) return INT32_MIN;
	dd = dr > dq ? dr - dq : dq - dr;
	if (dd > bw) return INT32_MIN;
	dg = dr < dq ? dr : dq;
	sc = ai->len < dg ? ai->len : dg;
	if (dd) {
		const float lin = pen_gap * (float)dd;
		const float lg = orc_log2((float)(dd + 1));
		const float half = .5f * lg;
		const float tot = lin + half;
		sc -= (int)tot;
	}
	return (int32_t)sc;
}

/* mb_chain_bk_end (lchain.c:25-42): where the backtrack from z[k] stops: at a visited anchor, or where
 * the chain's score since the best cut point has dropped by more than max_drop */
static int64_t chain_bk_end(int32_t max_drop, const orc_u128_t *z, const int32_t *f, const int64_t *p, int32_t *t, int64_t k)
{
	int64_t i = (int64_t)z[k].y, end_i = -1, max_i = i;
	int32_t max_s = 0;
	if (i < 0 || t[i] != 0) return i;
	do {
		int32_t s;
		t[i] = 2;
		end_i = i = p[i];
		s = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
		if (s > max_s) max_s = s, max_i = i;
		else if (max_s - s > max_drop) break;
	} while (i >= 0 && t[i] == 0);
	for (i = (int64_t)z[k].y; i >= 0 && i != end_i; i = p[i]) t[i] = 0;
	return max_i;
}

/* mb_lchain_dp without its allocator: a[0..n) in mb_anchor's order (edited); chains to u (score<<32 |
 * anchor count, n_u of them, in the order of their first anchor's concatenated tpos) and their anchors,
 * first to last, to b.  Returns n_u (0: no chain; -1: out of memory).  *n_b receives the anchor count. */
int32_t orc_lchain(int64_t n_ctg, const uint64_t *ctg_off, const uint64_t *ctg_len, int32_t max_dist_x, int32_t max_dist_y,
                   int32_t bw, int32_t max_skip, int32_t max_iter, int32_t min_sc, float pen_gap, int64_t n, const orc_anchor_t *a,
                   uint64_t *u_out, orc_anchor_t *b_out, int64_t *n_b)
{
	int32_t *f, *t, *v, n_u = 0, max_drop = bw;
	int64_t *p, i, j, k, max_ii, n_z, n_v = 0;
	orc_u128_t *z, *w;
	uint64_t *u;
	(void)n_ctg;
	*n_b = 0;
	if (n <= 0) return 0;
	if (max_dist_x < bw) max_dist_x = bw;
	if (max_dist_y < bw) max_dist_y = bw;
	f = (int32_t*)malloc(n * 4), t = (int32_t*)calloc(n, 4), v = (int32_t*)malloc(n * 4), p = (int64_t*)malloc(n * 8);
	z = (orc_u128_t*)malloc(n * sizeof(orc_u128_t)), u = (uint64_t*)malloc(n * 8);
	if (!f || !t || !v || !p || !z || !u) return -1;
	/* the DP (lchain.c:170-222) */
	for (i = 0, max_ii = -1; i < n; ++i) {
		int64_t max_j = -1, end_j;
		int32_t max_f = a[i].len, n_skip = 0;
		for (j = i - 1; j >= 0 && j >= i - max_iter; --j) {
			int32_t sc;
			if (a[i].tpos - a[j].tpos >= (int64_t)max_dist_x + a[i].len) break;
			sc = chain_sc(&a[i], &a[j], max_dist_x, max_dist_y, bw, pen_gap);
			if (sc == INT32_MIN) continue;
			sc += f[j];
			if (sc > max_f) {
				max_f = sc, max_j = j;
				if (n_skip > 0) --n_skip;
			} else if (t[j] == (int32_t)i) {
				if (++n_skip > max_skip) break;
			}
			if (p[j] >= 0) t[p[j]] = (int32_t)i;
		}
		end_j = j;
		if (max_ii < 0 || a[i].tpos - a[max_ii].tpos > (int64_t)max_dist_x + a[i].len) {
			int32_t mx = INT32_MIN;
			max_ii = -1;
			for (j = i - 1; j >= end_j && j >= 0; --j)
				if (mx < f[j]) mx = f[j], max_ii = j;
		}
		if (max_ii >= 0 && max_ii < end_j) {
			const int32_t tmp = chain_sc(&a[i], &a[max_ii], max_dist_x, max_dist_y, bw, pen_gap);
			if (tmp != INT32_MIN && max_f < tmp + f[max_ii]) max_f = tmp + f[max_ii], max_j = max_ii;
		}
		f[i] = max_f, p[i] = max_j;
		v[i] = max_j >= 0 && v[max_j] > max_f ? v[max_j] : max_f;
		if (max_ii < 0 || (a[i].tpos - a[max_ii].tpos <= (int64_t)max_dist_x + a[i].len && f[max_ii] < f[i])) max_ii = i;
	}
	/* the backtrack (lchain.c:44-92): ends by score, klib's radix order among equal scores; one pass
	 * (the reference counts first, then fills) */
	for (i = 0, n_z = 0; i < n; ++i)
		if (f[i] >= min_sc) z[n_z].x = (uint64_t)(int64_t)f[i], z[n_z++].y = (uint64_t)i;
	if (n_z > 0) {
		rs_sort(z, sizeof(*z), n_z, key_u128x);
		memset(t, 0, n * 4);
		for (k = n_z - 1; k >= 0; --k) {
			if (t[z[k].y] == 0) {
				const int64_t n_v0 = n_v, end_i = chain_bk_end(max_drop, z, f, p, t, k);
				int32_t sc;
				for (i = (int64_t)z[k].y; i != end_i; i = p[i]) v[n_v++] = (int32_t)i, t[i] = 1;
				sc = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
				if (sc >= min_sc && n_v > n_v0) u[n_u++] = (uint64_t)sc << 32 | (uint64_t)(n_v - n_v0);
				else n_v = n_v0;
			}
		}
	}
	/* compact_a (lchain.c:94-123): chains first anchor first, then ordered by the first anchor's
	 * concatenated tpos (klib's radix order among equal ones) */
	if (n_u > 0) {
		orc_anchor_t *bb = (orc_anchor_t*)malloc(n_v * sizeof(orc_anchor_t));
		w = (orc_u128_t*)malloc(n_u * sizeof(orc_u128_t));
		if (!bb || !w) return -1;
		for (i = 0, k = 0; i < n_u; ++i) {
			const int64_t k0 = k, ni = (int32_t)u[i];
			for (j = 0; j < ni; ++j) bb[k++] = a[v[k0 + (ni - j - 1)]];
		}
		for (i = 0, k = 0; i < n_u; ++i) {
			const int64_t c = bb[k].sid >> 1;
			w[i].x = (uint64_t)(bb[k].tpos + (int64_t)(ctg_off[c] * 2 + ctg_len[c] * (uint64_t)(bb[k].sid & 1)));
			w[i].y = (uint64_t)k << 32 | (uint64_t)i;
			k += (int32_t)u[i];
		}
		rs_sort(w, sizeof(*w), n_u, key_u128x);
		for (i = 0, k = 0; i < n_u; ++i) {
			const int32_t jj = (int32_t)w[i].y, m = (int32_t)u[jj];
			u_out[i] = u[jj];
			memcpy(&b_out[k], &bb[w[i].y >> 32], m * sizeof(orc_anchor_t));
			k += m;
		}
		*n_b = k;
		free(bb); free(w);
	}
	free(f); free(t); free(v); free(p); free(z); free(u);
	return n_u;
}

/* main.c:164-200 (`minibwa bench`): checksum over n splitmix64-drawn queries; type 0 = rank2a,
 * 1 = sa, 2 = batched sa over [k, k+intv) */
static inline uint64_t orc_sm64(uint64_t *x)
{
	uint64_t z = ((*x) += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

uint64_t orc_bench(const orc_bwt_t *b, int type, int64_t n, int intv)
{
	uint64_t x = 11, cs = 1;
	int64_t i;
	for (i = 0; i < n; ++i) {
		if (type == 0) {
			uint64_t k = orc_sm64(&x) % b->seq_len, l = orc_sm64(&x) % b->seq_len, ck[4], cl[4];
			orc_rank2a(b, k, l, ck, cl);
			cs = cs * ck[1] + cl[0];
		} else if (type == 1) {
			uint64_t k = orc_sm64(&x) % b->seq_len;
			cs = cs * 0xbf58476d1ce4e5b9ULL ^ orc_sa(b, k);
		} else {
			uint64_t j, xr = 0, k = orc_sm64(&x) % b->seq_len;
			uint64_t l = k + intv < b->seq_len ? k + intv : b->seq_len;
			for (j = k; j < l; ++j) xr ^= orc_sa(b, j);
			cs = cs * 0xbf58476d1ce4e5b9ULL ^ xr;
		}
	}
	return cs;
}
