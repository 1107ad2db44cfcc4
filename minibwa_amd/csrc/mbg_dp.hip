// mbg_dp.hip -- banded dual-affine (ksw_extd2) and single-affine (ksw_extz2) extension DP for gfx950.
//
// One 64-lane wavefront per job; each anti-diagonal r is computed by the lanes in parallel and the
// t-1 neighbour of the previous row arrives by a DPP lane shift.  The results must equal
// ksw_extd2_sse (ksw2_extd2_sse.c:27-403) and ksw_extz2_sse (ksw2_extz2_sse.c:16-296) bit for bit,
// so the kernel keeps their observable semantics rather than a textbook recurrence:
//   * int8 difference arithmetic with wrap-around, and the unsigned/signed max mix of extz2;
//   * every row covers the 16-aligned range [st0/16*16, (en0+16)/16*16-1]: cells outside the exact
//     band are computed, kept and later read, exactly as the 16-lane SSE blocks do;
//   * the score row is written in 16-cell groups starting at st0 (so it runs past en0 with the
//     target's zero padding) and cells it does not reach keep their stale scores;
//   * an index of the row state takes the reference's initial value when it first enters the
//     window [.., en+15] (memset/calloc/NEG_INF of the reference's arrays);
//   * the exact-max argmax reproduces the SSE reduction order (H[en0] first, four lanes with
//     t = st0 mod 4 classes keeping their first strict maximum, merged in lane order, then the
//     scalar tail), encoded as one priority key per cell and wave max-reductions;
//   * the backtrack walks the rotated matrix (ksw2.h:140-172) from LDS-staged chunks of rows.
//
#include "mbg_internal.h"
#include "mbg_wave.h"

namespace mbg {

constexpr int32_t kNegInf = -0x40000000;

__device__ __forceinline__ int8_t smax8(int8_t a, int8_t b) { return a > b ? a : b; }
__device__ __forceinline__ int8_t smin8(int8_t a, int8_t b) { return a < b ? a : b; }

struct DpState { int32_t max, zdropped, max_q, max_t, mqe, mqe_t, mte, mte_q, score, reach_end; };

// ksw_apply_zdrop (ksw2.h:181-197), rotated coordinates; written as selects so the state stays in
// registers
__device__ __forceinline__ bool apply_zdrop(DpState &z, int32_t H, int r, int t, int zdrop, int e)
{
	const bool up = H > z.max;
	const int tl = t - z.max_t, ql = (r - t) - z.max_q, l = tl > ql ? tl - ql : ql - tl;
	const bool drop = !up && t >= z.max_t && r - t >= z.max_q && zdrop >= 0 && z.max - H > zdrop + l * e;
	z.max = up ? (H & 0x7fffffff) : z.max;
	z.max_t = up ? t : z.max_t;
	z.max_q = up ? r - t : z.max_q;
	z.zdropped = drop ? 1 : z.zdropped;
	return drop;
}

// per-job parameters (ksw2_extd2_sse.c:60-97 / ksw2_extz2_sse.c:41-74)
template <bool DUAL>
struct JobP {
	int qlen, tlen, m, flag, zdrop, end_bonus;
	int8_t q, e, q2, e2;                 // after extd2's swap that makes (q,e) the cheaper opening
	int qe0;                             // extd2's scalar q+e, taken before the swap
	int qe;                              // extz2's scalar q+e
	int8_t qe8, qe28, sc_mch, sc_mis, m1, sc_N, qe2x, max_sc_, ini, ini2;
	int w, n_col_, nb, tl_hi, max_sc, long_thres, long_diff;
	bool skip, with_cigar, approx_max, right, generic;
	bool unplanned;                      // outside the launch's plan: not run, reported with cells -1
};

// the scoring matrix enters through scalars the host derived from it (entries 0, 1 and m*m-1, the
// max and the min), so that every job parameter stays wave-uniform
template <bool DUAL>
__device__ __forceinline__ JobP<DUAL> job_params(const DpJob &jb, const DpArgs &a)
{
	const int m = a.m;
	JobP<DUAL> P;
	P.qlen = jb.qlen, P.tlen = jb.tlen, P.m = m, P.flag = jb.flag, P.zdrop = jb.zdrop, P.end_bonus = jb.end_bonus;
	int8_t q = jb.q, e = jb.e, q2 = jb.q2, e2 = jb.e2;
	P.with_cigar = !(P.flag & 0x01), P.approx_max = !!(P.flag & 0x08), P.right = !!(P.flag & 0x02);
	P.generic = !!(P.flag & 0x04);
	P.skip = DUAL ? (m <= 1 || P.qlen <= 0 || P.tlen <= 0) : (m <= 0 || P.qlen <= 0 || P.tlen <= 0);
	P.qe0 = q + e;
	if (DUAL && q2 + e2 < q + e) { int8_t t = q; q = q2; q2 = t; t = e; e = e2; e2 = t; }
	P.q = q, P.e = e, P.q2 = q2, P.e2 = e2;
	P.qe8 = (int8_t)(q + e), P.qe28 = (int8_t)(q2 + e2);
	P.qe = q + e;
	P.sc_mch = a.mat[0], P.sc_mis = a.mat[1], P.m1 = (int8_t)(m - 1);
	P.sc_N = 0, P.qe2x = 0, P.max_sc_ = 0;
	P.w = jb.w, P.n_col_ = 1, P.max_sc = 0, P.long_thres = 0, P.long_diff = 0;
	if (!P.skip) {
		P.sc_N = DUAL ? (a.mat_last == 0 ? (int8_t)-e2 : (int8_t)a.mat_last) : (a.mat_last == 0 ? (int8_t)-e : (int8_t)a.mat_last);
		P.qe2x = (int8_t)((q + e) * 2);
		P.max_sc_ = (int8_t)(a.mat[0] + (q + e) * 2);
		dp_geom(P.qlen, P.tlen, P.w, &P.w, &P.n_col_);
		P.max_sc = a.mat_max;
		if (-a.mat_min > 2 * (q + e)) P.skip = true;
		if (DUAL) {
			P.long_thres = e != e2 ? (q2 - q) / (e - e2) - 1 : 0;
			if (q2 + e2 + P.long_thres * e2 > q + e + P.long_thres * e) ++P.long_thres;
			P.long_diff = P.long_thres * (e - e2) - (q2 - q) - e2;
		}
	}
	P.nb = P.n_col_ * 16;
	// a job the launch was not planned for (the *_dev entry points size slabs and rings from the
	// caller's max_qlen / max_tlen / max_w) would overrun its block's backtrack slab or find no ring
	// wide enough: it runs as an empty job and reports cells -1 instead
	P.unplanned = false;
	if (!P.skip) {
		const bool ring = P.nb + 17 <= a.cap || (a.gcap > 0 && P.nb + 17 <= a.gcap) ||
		                  (a.tcap > 0 && P.with_cigar && !P.approx_max && !P.generic && P.nb + 144 <= a.tcap && P.qlen + P.tlen <= 30000);
		if ((int64_t)(P.qlen + P.tlen) * P.nb + 256 > a.slab || !ring) P.unplanned = P.skip = true, P.n_col_ = 1, P.nb = 16;
	}
	P.ini = DUAL ? (int8_t)(-q - e) : 0, P.ini2 = DUAL ? (int8_t)(-q2 - e2) : 0;
	P.tl_hi = (P.tlen + 15) / 16 * 16 - 1;
	return P;
}

// band of anti-diagonal r (ksw2_extd2_sse.c:130-139); false when it is empty
__device__ __forceinline__ bool row_range(int r, int qlen, int tlen, int w, int &st0, int &en0)
{
	int st = 0, en = tlen - 1;
	if (st < r - qlen + 1) st = r - qlen + 1;
	if (en > r) en = r;
	if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
	if (en > (r + w) >> 1) en = (r + w) >> 1;
	st0 = st, en0 = en;
	return st <= en;
}

// values of cell st-1 when the previous row did not hold it (ksw2_extd2_sse.c:141-155 /
// ksw2_extz2_sse.c:110-115)
template <bool DUAL>
__device__ __forceinline__ void row_boundary(const JobP<DUAL> &P, int r, int st, int8_t &x1, int8_t &v1, int8_t &x21)
{
	if (DUAL) {
		x1 = (int8_t)(-P.q - P.e), x21 = (int8_t)(-P.q2 - P.e2);
		if (st > 0) v1 = (int8_t)(-P.q - P.e);
		else v1 = (int8_t)(r == 0 ? -P.q - P.e : r < P.long_thres ? -P.e : r == P.long_thres ? P.long_diff : -P.e2);
	} else {
		x1 = 0, x21 = 0;
		v1 = st > 0 ? 0 : (int8_t)(r ? P.q : 0);
	}
}

// the values index r takes when the row reaches it (ksw2_extd2_sse.c:156-161)
template <bool DUAL>
__device__ __forceinline__ void row_yu(const JobP<DUAL> &P, int r, int8_t &y, int8_t &u, int8_t &y2)
{
	if (DUAL) {
		y = (int8_t)(-P.q - P.e), y2 = (int8_t)(-P.q2 - P.e2);
		u = (int8_t)(r == 0 ? -P.q - P.e : r < P.long_thres ? -P.e : r == P.long_thres ? P.long_diff : -P.e2);
	} else {
		y = 0, y2 = 0, u = (int8_t)(r ? P.q : 0);
	}
}

// one cell of the difference recurrence (ksw2_extd2_sse.c:186-265 / ksw2_extz2_sse.c:139-229);
// returns the backtrack byte
template <bool DUAL>
__device__ __forceinline__ uint8_t dp_cell(const JobP<DUAL> &P, int8_t xt1, int8_t vt1, int8_t x2t1, int8_t ou, int8_t oy,
                                           int8_t oy2, int8_t zs, int8_t &nU, int8_t &nV, int8_t &nX, int8_t &nY,
                                           int8_t &nX2, int8_t &nY2)
{
	uint8_t d = 0;
	if (DUAL) {
		int8_t zz = zs;
		int8_t A = (int8_t)(xt1 + vt1), B = (int8_t)(oy + ou), A2 = (int8_t)(x2t1 + vt1), B2 = (int8_t)(oy2 + ou);
		if (!P.with_cigar) {
			zz = smax8(zz, A); zz = smax8(zz, B); zz = smax8(zz, A2); zz = smax8(zz, B2);
		} else if (!P.right) {
			d = A > zz ? 1 : 0;   zz = smax8(zz, A);
			d = B > zz ? 2 : d;   zz = smax8(zz, B);
			d = A2 > zz ? 3 : d;  zz = smax8(zz, A2);
			d = B2 > zz ? 4 : d;  zz = smax8(zz, B2);
		} else {
			d = zz > A ? 0 : 1;   zz = smax8(zz, A);
			d = zz > B ? d : 2;   zz = smax8(zz, B);
			d = zz > A2 ? d : 3;  zz = smax8(zz, A2);
			d = zz > B2 ? d : 4;  zz = smax8(zz, B2);
		}
		zz = smin8(zz, P.sc_mch);
		nU = (int8_t)(zz - vt1);
		nV = (int8_t)(zz - ou);
		int8_t tmp = (int8_t)(zz - P.q);
		A = (int8_t)(A - tmp); B = (int8_t)(B - tmp);
		tmp = (int8_t)(zz - P.q2);
		A2 = (int8_t)(A2 - tmp); B2 = (int8_t)(B2 - tmp);
		if (!P.with_cigar || !P.right) {
			nX  = (int8_t)((A > 0 ? A : 0) - P.qe8);
			nY  = (int8_t)((B > 0 ? B : 0) - P.qe8);
			nX2 = (int8_t)((A2 > 0 ? A2 : 0) - P.qe28);
			nY2 = (int8_t)((B2 > 0 ? B2 : 0) - P.qe28);
			d |= (A > 0 ? 0x08 : 0) | (B > 0 ? 0x10 : 0) | (A2 > 0 ? 0x20 : 0) | (B2 > 0 ? 0x40 : 0);
		} else {
			nX  = (int8_t)((A >= 0 ? A : 0) - P.qe8);
			nY  = (int8_t)((B >= 0 ? B : 0) - P.qe8);
			nX2 = (int8_t)((A2 >= 0 ? A2 : 0) - P.qe28);
			nY2 = (int8_t)((B2 >= 0 ? B2 : 0) - P.qe28);
			d |= (A >= 0 ? 0x08 : 0) | (B >= 0 ? 0x10 : 0) | (A2 >= 0 ? 0x20 : 0) | (B2 >= 0 ? 0x40 : 0);
		}
	} else {
		const uint8_t uvt1 = (uint8_t)vt1, uut = (uint8_t)ou;
		int8_t zz = (int8_t)((uint8_t)zs + (uint8_t)P.qe2x);
		int8_t A = (int8_t)((uint8_t)xt1 + uvt1), B = (int8_t)((uint8_t)oy + uut);
		if (!P.with_cigar) {
			zz = smax8(zz, A);
		} else if (!P.right) {
			d = A > zz ? 1 : 0;
			zz = smax8(zz, A);
			d = B > zz ? 2 : d;
		} else {
			d = zz > A ? 0 : 1;
			zz = smax8(zz, A);
			d = zz > B ? d : 2;
		}
		uint8_t zu = (uint8_t)zz > (uint8_t)B ? (uint8_t)zz : (uint8_t)B;
		zu = zu < (uint8_t)P.max_sc_ ? zu : (uint8_t)P.max_sc_;
		nU = (int8_t)(uint8_t)(zu - uvt1);
		nV = (int8_t)(uint8_t)(zu - uut);
		const int8_t zq = (int8_t)(uint8_t)(zu - (uint8_t)P.q);
		A = (int8_t)(A - zq); B = (int8_t)(B - zq);
		nX2 = 0, nY2 = 0;
		if (!P.with_cigar || !P.right) {
			nX = A > 0 ? A : 0; nY = B > 0 ? B : 0;
			d |= (A > 0 ? 0x08 : 0) | (B > 0 ? 0x10 : 0);
		} else {
			nX = A >= 0 ? A : 0; nY = B >= 0 ? B : 0;
			d |= (A >= 0 ? 0x08 : 0) | (B >= 0 ? 0x10 : 0);
		}
	}
	return d;
}

// the cell of dp_cell for jobs with a CIGAR, the alignment side fixed at compile time (RIGHT: the
// KSW_EZ_RIGHT tie rule, ksw2_extd2_sse.c:233-265) and both branches written as selects
template <bool DUAL, bool RIGHT>
__device__ __forceinline__ uint8_t dp_cell_cg(const JobP<DUAL> &P, int8_t xt1, int8_t vt1, int8_t x2t1, int8_t ou, int8_t oy,
                                              int8_t oy2, int8_t zs, int8_t &nU, int8_t &nV, int8_t &nX, int8_t &nY,
                                              int8_t &nX2, int8_t &nY2)
{
	uint8_t d;
	if (DUAL) {
		int8_t zz = zs;
		int8_t A = (int8_t)(xt1 + vt1), B = (int8_t)(oy + ou), A2 = (int8_t)(x2t1 + vt1), B2 = (int8_t)(oy2 + ou);
		if (!RIGHT) {
			d = A > zz ? 1 : 0;   zz = smax8(zz, A);
			d = B > zz ? 2 : d;   zz = smax8(zz, B);
			d = A2 > zz ? 3 : d;  zz = smax8(zz, A2);
			d = B2 > zz ? 4 : d;  zz = smax8(zz, B2);
		} else {
			d = zz > A ? 0 : 1;   zz = smax8(zz, A);
			d = zz > B ? d : 2;   zz = smax8(zz, B);
			d = zz > A2 ? d : 3;  zz = smax8(zz, A2);
			d = zz > B2 ? d : 4;  zz = smax8(zz, B2);
		}
		zz = smin8(zz, P.sc_mch);
		nU = (int8_t)(zz - vt1);
		nV = (int8_t)(zz - ou);
		int8_t tmp = (int8_t)(zz - P.q);
		A = (int8_t)(A - tmp); B = (int8_t)(B - tmp);
		tmp = (int8_t)(zz - P.q2);
		A2 = (int8_t)(A2 - tmp); B2 = (int8_t)(B2 - tmp);
		const bool pA = RIGHT ? A >= 0 : A > 0, pB = RIGHT ? B >= 0 : B > 0;
		const bool pA2 = RIGHT ? A2 >= 0 : A2 > 0, pB2 = RIGHT ? B2 >= 0 : B2 > 0;
		nX  = (int8_t)((pA ? A : 0) - P.qe8);
		nY  = (int8_t)((pB ? B : 0) - P.qe8);
		nX2 = (int8_t)((pA2 ? A2 : 0) - P.qe28);
		nY2 = (int8_t)((pB2 ? B2 : 0) - P.qe28);
		d |= (pA ? 0x08 : 0) | (pB ? 0x10 : 0) | (pA2 ? 0x20 : 0) | (pB2 ? 0x40 : 0);
	} else {
		const uint8_t uvt1 = (uint8_t)vt1, uut = (uint8_t)ou;
		int8_t zz = (int8_t)((uint8_t)zs + (uint8_t)P.qe2x);
		int8_t A = (int8_t)((uint8_t)xt1 + uvt1), B = (int8_t)((uint8_t)oy + uut);
		if (!RIGHT) {
			d = A > zz ? 1 : 0;
			zz = smax8(zz, A);
			d = B > zz ? 2 : d;
		} else {
			d = zz > A ? 0 : 1;
			zz = smax8(zz, A);
			d = zz > B ? d : 2;
		}
		uint8_t zu = (uint8_t)zz > (uint8_t)B ? (uint8_t)zz : (uint8_t)B;
		zu = zu < (uint8_t)P.max_sc_ ? zu : (uint8_t)P.max_sc_;
		nU = (int8_t)(uint8_t)(zu - uvt1);
		nV = (int8_t)(uint8_t)(zu - uut);
		const int8_t zq = (int8_t)(uint8_t)(zu - (uint8_t)P.q);
		A = (int8_t)(A - zq); B = (int8_t)(B - zq);
		nX2 = 0, nY2 = 0;
		const bool pA = RIGHT ? A >= 0 : A > 0, pB = RIGHT ? B >= 0 : B > 0;
		nX = pA ? A : 0; nY = pB ? B : 0;
		d |= (pA ? 0x08 : 0) | (pB ? 0x10 : 0);
	}
	return d;
}

// H increment carried by a new v (cells t < en0) or u (cell en0) value
template <bool DUAL>
__device__ __forceinline__ int32_t h_term(const JobP<DUAL> &P, int8_t x)
{
	return DUAL ? (int32_t)x : (int32_t)(uint8_t)x - P.qe;
}

// priority of cell t in the SSE argmax order: cells of the 4-lane body rank by lane class
// (t - st0) % 4 first, then by position; the scalar tail ranks below them, earliest first; H[en0]
// (key 0xffffffff) ranks above everything
__device__ __forceinline__ uint32_t argmax_key(int t, int st0, int en1)
{
	return t < en1 ? (0x80000000u | ((uint32_t)(3 - ((t - st0) & 3)) << 28) | (0x0FFFFFFFu - (uint32_t)((t - st0) >> 2)))
	               : (0x7FFFFFFFu - (uint32_t)(t - en1));
}

// the same order packed with the score for rows of at most 247 cells and |H| < 2^23:
// (H + 2^23) << 8 | rank, rank 255 for en0, 254 - (class*63 + position) for body cells, 2 - offset
// for the (at most 3) tail cells
__device__ __forceinline__ uint32_t argmax_key8(int32_t h, int t, int st0, int en1)
{
	const int d = t - st0;
	const uint32_t rank = t < en1 ? (uint32_t)(254 - ((d & 3) * 63 + (d >> 2))) : (uint32_t)(2 - (t - en1));
	return ((uint32_t)(h + (1 << 23)) << 8) | rank;
}

__device__ __forceinline__ void argmax_decode8(uint32_t k, int st0, int en0, int en1, int32_t &max_H, int &max_t)
{
	max_H = (int32_t)(k >> 8) - (1 << 23);
	const int rank = (int)(k & 255u);
	if (rank == 255) max_t = en0;
	else if (rank >= 3) { const int v = 254 - rank; max_t = st0 + (v % 63) * 4 + v / 63; }
	else max_t = en1 + (2 - rank);
}

__device__ __forceinline__ void argmax_take(uint32_t &bh, uint32_t &bp, int32_t h, uint32_t key)
{
	const uint32_t hk = (uint32_t)h ^ 0x80000000u;
	if (hk > bh || (hk == bh && key > bp)) bh = hk, bp = key;
}

__device__ __forceinline__ void argmax_reduce(uint32_t bh, uint32_t bp, int st0, int en0, int en1, int32_t &max_H, int &max_t)
{
	const uint32_t mh = wave_max_u32(bh);
	const uint32_t pri = wave_max_u32(bh == mh ? bp : 0u);
	max_H = (int32_t)(mh ^ 0x80000000u);
	if (pri == 0xFFFFFFFFu) max_t = en0;
	else if (pri & 0x80000000u) max_t = st0 + (int)(0x0FFFFFFFu - (pri & 0x0FFFFFFFu)) * 4 + (3 - (int)((pri >> 28) & 3));
	else max_t = en1 + (int)(0x7FFFFFFFu - pri);
}

// end of an exact-max row (ksw2_extd2_sse.c:345-367); true when the DP stops
template <bool DUAL>
__device__ __forceinline__ bool finish_exact(const JobP<DUAL> &P, DpState &z, int r, int st0, int en0, int32_t max_H, int max_t,
                                             int32_t Hen0, int32_t Hst0, int32_t &last_max_H)
{
	if (en0 == P.tlen - 1 && Hen0 > z.mte) z.mte = Hen0, z.mte_q = r - en0;
	if (r - st0 == P.qlen - 1 && Hst0 > z.mqe) z.mqe = Hst0, z.mqe_t = st0;
	if (apply_zdrop(z, max_H, r, max_t, P.zdrop, DUAL ? P.e2 : P.e)) return true;
	bool stop = false;
	if (r == P.qlen + P.tlen - 2 && en0 == P.tlen - 1) z.score = Hen0;
	if (DUAL && (P.flag & 0x40)) { // ksw2_extd2_sse.c:361-367
		const int32_t rH = last_max_H > max_H ? last_max_H : max_H;
		const int32_t rq = P.qlen - (r - st0), rt = P.tlen - en0;
		const int32_t rm = rq >= P.tlen - st0 ? P.tlen - st0 : rt >= P.qlen - (r - en0) ? P.qlen - (r - en0) : P.tlen + P.qlen - 1 - r;
		if (rH + rm * P.max_sc + P.end_bonus < z.max) stop = true;
	}
	last_max_H = max_H;
	return stop;
}

// approximate max (ksw2_extd2_sse.c:369-385 / ksw2_extz2_sse.c:262-278): H0 follows one path
// along the rows; rdV(t)/rdU(t) read the row's new v/u at index t
template <bool DUAL, class RV, class RU>
__device__ __forceinline__ bool finish_approx(const JobP<DUAL> &P, DpState &z, int r, int st0, int en0, int32_t &H0, int &last_H0_t,
                                              RV &&rdV, RU &&rdU)
{
	bool stop = false;
	if (r > 0) {
		const bool in0 = last_H0_t >= st0 && last_H0_t <= en0, in1 = last_H0_t + 1 >= st0 && last_H0_t + 1 <= en0;
		if (in0 && in1) {
			const int32_t d0 = h_term<DUAL>(P, rdV(last_H0_t));
			const int32_t d1 = h_term<DUAL>(P, rdU(last_H0_t + 1));
			if (d0 > d1) H0 += d0;
			else H0 += d1, ++last_H0_t;
		} else if (in0) {
			H0 += h_term<DUAL>(P, rdV(last_H0_t));
		} else {
			++last_H0_t;
			H0 += h_term<DUAL>(P, rdU(last_H0_t));
		}
		if (!DUAL && (P.flag & 0x10) && apply_zdrop(z, H0, r, last_H0_t, P.zdrop, P.e)) stop = true;
	} else {
		const int8_t v0 = rdV(0);
		H0 = DUAL ? (int32_t)v0 - P.qe0 : (int32_t)(uint8_t)v0 - P.qe - P.qe;
		last_H0_t = 0;
	}
	if (!stop && DUAL && (P.flag & 0x10) && apply_zdrop(z, H0, r, last_H0_t, P.zdrop, P.e2)) stop = true;
	if (!stop && r == P.qlen + P.tlen - 2 && en0 == P.tlen - 1) z.score = H0;
	return stop;
}

// backtrack (ksw2_extd2_sse.c:391-402, ksw2.h:140-172) over the rows kept in `slab`, staged through
// the first `stage` bytes of LDS, or read from the slab itself (DIRECT), by a workgroup of NT
// threads that all walk the same path; returns n_cigar
template <bool DUAL, bool DIRECT = false, int NT = 64>
__device__ __forceinline__ int backtrack(const JobP<DUAL> &P, DpState &z, const DpJob &jb, const DpArgs &a, uint8_t *lds, const uint8_t *slab,
                                         int stage)
{
	const int lane = threadIdx.x;
	const int qlen = P.qlen, tlen = P.tlen, w = P.w, nb = P.nb;
	uint32_t *cg = a.cigar + jb.cigar_off;
	int n_cigar = 0;
	int i0 = -1, j0 = -1;
	if (!P.skip && P.with_cigar) {
		if (!z.zdropped && !(P.flag & 0x40)) i0 = tlen - 1, j0 = qlen - 1;
		else if ((P.flag & 0x40) && z.mqe + P.end_bonus > z.max) z.reach_end = 1, i0 = z.mqe_t, j0 = qlen - 1;
		else if (z.max_t >= 0 && z.max_q >= 0) i0 = z.max_t, j0 = z.max_q;
	}
	if (i0 < 0 && j0 < 0) return 0;
	uint8_t *stg = lds;
	const int K = stage / nb;
	int chunk_lo = 0x7fffffff;
	int i = i0, j = j0, state = 0, last_op = -1;
	uint32_t cur = 0;
	while (i >= 0 && j >= 0) {
		const int r = i + j;
		int s0, e0;
		row_range(r, qlen, tlen, w, s0, e0);
		const int off = s0 >> 4 << 4, off_end = ((e0 + 16) >> 4 << 4) - 1;
		int force = -1;
		if (i < off) force = 2;
		if (i > off_end) force = 1;
		uint32_t tmp = 0;
		if (force < 0 && DIRECT) {
			tmp = (uint32_t)__builtin_amdgcn_readfirstlane(slab[(size_t)r * nb + i - off]);
		} else if (force < 0) {
			if (r < chunk_lo) {
				const int lo = r - K + 1 > 0 ? r - K + 1 : 0;
				const int bytes = (r - lo + 1) * nb;
				__syncthreads();
				for (int o = lane * 16; o < bytes; o += NT * 16)
					*(uint4*)(stg + o) = *(const uint4*)(slab + (size_t)lo * nb + o);
				__syncthreads();
				chunk_lo = lo;
			}
			// every lane reads the same byte: marked uniform, so the walk (and the barriers of the
			// staging above) stays in uniform control flow
			tmp = (uint32_t)__builtin_amdgcn_readfirstlane(stg[(r - chunk_lo) * nb + i - off]);
		}
		if (state == 0) state = tmp & 7;
		else if (!(tmp >> (state + 2) & 1)) state = 0;
		if (state == 0) state = tmp & 7;
		if (force >= 0) state = force;
		int op;
		if (state == 0) op = 0, --i, --j;
		else if (state == 1 || state == 3) op = 2, --i;
		else op = 1, --j;
		if (op == last_op) cur += 16;
		else {
			if (last_op >= 0 && lane == 0) cg[n_cigar - 1] = cur;
			++n_cigar, last_op = op, cur = 16u | (uint32_t)op;
		}
	}
	if (i >= 0) {
		if (last_op == 2) cur += (uint32_t)(i + 1) << 4;
		else { if (last_op >= 0 && lane == 0) cg[n_cigar - 1] = cur; ++n_cigar, last_op = 2, cur = (uint32_t)(i + 1) << 4 | 2; }
	}
	if (j >= 0) {
		if (last_op == 1) cur += (uint32_t)(j + 1) << 4;
		else { if (last_op >= 0 && lane == 0) cg[n_cigar - 1] = cur; ++n_cigar, last_op = 1, cur = (uint32_t)(j + 1) << 4 | 1; }
	}
	if (last_op >= 0 && lane == 0) cg[n_cigar - 1] = cur;
	if (!(P.flag & 0x80)) {
		__syncthreads();
		for (int k = lane; k < n_cigar / 2; k += NT) {
			const uint32_t t0 = cg[k], t1 = cg[n_cigar - 1 - k];
			cg[k] = t1, cg[n_cigar - 1 - k] = t0;
		}
	}
	return n_cigar;
}

__device__ __forceinline__ void init_state(DpState &z)
{
	z.max_q = z.max_t = z.mqe_t = z.mte_q = -1;
	z.max = 0, z.score = z.mqe = z.mte = kNegInf;
	z.zdropped = 0, z.reach_end = 0;
}

__device__ __forceinline__ void write_result(const DpArgs &a, int jid, const DpState &z, int n_cigar, int cells)
{
	if (threadIdx.x == 0) {
		DpRes o;
		o.max = z.max, o.zdropped = z.zdropped, o.max_q = z.max_q, o.max_t = z.max_t;
		o.mqe = z.mqe, o.mqe_t = z.mqe_t, o.mte = z.mte, o.mte_q = z.mte_q;
		o.score = z.score, o.reach_end = z.reach_end, o.n_cigar = n_cigar, o.cells = cells;
		a.res[jid] = o;
	}
}

// ---- the row loop --------------------------------------------------------------------------------
//
// One pass per 64-cell block of [st, hi], hi = min(en + 15, padded tlen - 1): indices entering the
// window take their initial values, the score row is refreshed on [st0, fresh_hi], the cells of
// [st, en] are computed and the exact-max H row and argmax keys updated, all with per-lane selects and
// one predicated store, so a row costs one trip per block instead of separate passes with barriers.
// A single wave owns the job, so its LDS reads see its earlier writes in program order.

// STAGED: the job's query and target are copied into LDS after the 12*cap state region first.
// GLOBAL: `lds` is the block's 12*CAP-byte state region in device memory, for bands wider than the
// LDS ring; a workgroup-scope fence ends each row so the next row's loads see its stores.
template <bool DUAL, bool STAGED, bool GLOBAL = false>
__device__ __forceinline__ void run_job_lds(const DpArgs &a, int jid, const DpJob &jb, const JobP<DUAL> P, uint8_t *lds, const int8_t *mat,
                                            uint8_t *slab, const int CAP)
{
	const int lane = threadIdx.x;
	const int CM = CAP - 1;
	int8_t *U = (int8_t*)lds, *V = U + CAP, *X = V + CAP, *Y = X + CAP, *X2 = Y + CAP, *Y2 = X2 + CAP, *S = Y2 + CAP;
	int32_t *H = (int32_t*)(lds + 8 * CAP);
	const uint8_t *Qg = a.qseq + jb.qoff, *Tg = a.tseq + jb.toff;
	const int qlen = P.qlen, tlen = P.tlen;
	uint8_t *Qs = lds + 12 * CAP, *Ts = Qs + ((qlen + 15) & ~15);
	if (STAGED && !P.skip) {
		for (int i = lane; i < qlen; i += 64) Qs[i] = Qg[i];
		for (int i = lane; i < tlen; i += 64) Ts[i] = Tg[i];
		__syncthreads();
	}
	const uint8_t *Q = STAGED ? Qs : Qg, *T = STAGED ? Ts : Tg;
	const bool hsmall = qlen + tlen <= 60000;   // |H| < 2^23: H and the argmax rank fit one 32-bit key
	DpState z;
	init_state(z);
	int cells = 0;
	int hw = -1, last_st = -1, last_en = -1, last_max_H = 0, last_H0_t = 0;
	int32_t H0 = 0;
	for (int r = 0; !P.skip && r < qlen + tlen - 1; ++r) {
		int st0, en0;
		if (!row_range(r, qlen, tlen, P.w, st0, en0)) { z.zdropped = 1; break; }
		cells += en0 - st0 + 1;
		const int st = st0 >> 4 << 4, en = ((en0 + 16) >> 4 << 4) - 1;
		int hi = en + 15;
		if (hi > P.tl_hi) hi = P.tl_hi;
		const int hw0 = hw;                          // indices above hw0 enter the window in this row
		if (hi > hw) hw = hi;
		// cell st-1 of the previous row, or the boundary constants (ksw2_extd2_sse.c:141-155)
		int8_t x1, v1, x21;
		if (st > 0 && st - 1 >= last_st && st - 1 <= last_en) {
			const int sl = (st - 1) & CM;
			x1 = X[sl], v1 = V[sl], x21 = DUAL ? X2[sl] : 0;
		} else row_boundary<DUAL>(P, r, st, x1, v1, x21);
		const int ep = en0 > 0 ? en0 - 1 : en0;
		const int32_t Hprev = ep > hw0 ? kNegInf : H[ep & CM];   // H[en0-1] before this row's update
		int8_t yr, ur, y2r;
		row_yu<DUAL>(P, r, yr, ur, y2r);
		const bool set_r = en >= r;
		int shi = st0 + ((en0 - st0) >> 4 << 4) + 15;
		if (shi > P.tl_hi) shi = P.tl_hi;
		const int fresh_hi = P.generic ? en0 : shi;  // the score row is rewritten on [st0, fresh_hi]
		const int en1 = st0 + ((en0 - st0) >> 2 << 2);
		const bool exact = !P.approx_max && r > 0;
		const bool packed = exact && hsmall && en0 - st0 <= 246;
		uint32_t bh = 0, bp = 0;                      // best key (packed) or (H, priority) of this lane
		int32_t h_en0 = 0, h_st0 = 0;
		uint8_t *prow = slab + (size_t)r * P.nb;
		int cx = x1, cv = v1, cx2 = x21;
		for (int base = st; base <= hi; base += 64) {
			const int t = base + lane, sl = t & CM;
			const bool entering = t > hw0;
			int8_t ox = X[sl], ov = V[sl], ou = U[sl], oy = Y[sl], zs = S[sl];
			int8_t ox2 = DUAL ? X2[sl] : 0, oy2 = DUAL ? Y2[sl] : 0;
			int32_t oh = H[sl];
			if (entering) {
				ox = ov = ou = oy = P.ini;
				ox2 = oy2 = P.ini2;
				zs = 0;
				oh = kNegInf;
			}
			if (set_r && t == r) { oy = yr; ou = ur; oy2 = y2r; }
			if (t >= st0 && t <= fresh_hi) {
				if (!P.generic) {
					const int8_t ta = t < tlen ? (int8_t)T[t] : 0, qa = t <= r ? (int8_t)Q[r - t] : 0;
					zs = (ta == P.m1 || qa == P.m1) ? P.sc_N : ta == qa ? P.sc_mch : P.sc_mis;
				} else {
					zs = mat[T[t] * P.m + Q[r - t]];
				}
			}
			int8_t xt1 = (int8_t)wave_shr1_fill((int)ox, cx), vt1 = (int8_t)wave_shr1_fill((int)ov, cv);
			int8_t x2t1 = DUAL ? (int8_t)wave_shr1_fill((int)ox2, cx2) : 0;
			if (!DUAL && base == st && lane >= 1 && lane <= 3) {
				// _mm_cvtsi32_si128 of a negative int8 also sets bytes 1..3 (ksw2_extz2_sse.c:138-139)
				if (x1 < 0) xt1 = (int8_t)0xff;
				if (v1 < 0) vt1 = (int8_t)0xff;
			}
			cx = __builtin_amdgcn_readlane((int)ox, 63);
			cv = __builtin_amdgcn_readlane((int)ov, 63);
			if (DUAL) cx2 = __builtin_amdgcn_readlane((int)ox2, 63);
			const bool core = t <= en;
			int8_t nU, nV, nX, nY, nX2, nY2;
			const uint8_t d = dp_cell<DUAL>(P, xt1, vt1, x2t1, ou, oy, oy2, zs, nU, nV, nX, nY, nX2, nY2);
			int32_t nh = oh;
			if (exact && t >= st0 && t <= en0) {
				nh = t < en0 ? oh + h_term<DUAL>(P, nV) : Hprev + h_term<DUAL>(P, en0 > 0 ? nU : nV);
				if (packed) {
					const uint32_t k = t < en0 ? argmax_key8(nh, t, st0, en1) : (((uint32_t)(nh + (1 << 23)) << 8) | 255u);
					bh = k > bh ? k : bh;
				} else {
					argmax_take(bh, bp, nh, t < en0 ? argmax_key(t, st0, en1) : 0xFFFFFFFFu);
				}
				if (t == en0) h_en0 = nh;
				if (t == st0) h_st0 = nh;
			}
			if (t <= hi) {
				U[sl] = core ? nU : ou; V[sl] = core ? nV : ov; X[sl] = core ? nX : ox; Y[sl] = core ? nY : oy;
				if (DUAL) { X2[sl] = core ? nX2 : ox2; Y2[sl] = core ? nY2 : oy2; }
				S[sl] = zs;
				H[sl] = nh;
				if (core && P.with_cigar) prow[t - st] = d;
			}
		}
		if (GLOBAL) __threadfence_block();
		__builtin_amdgcn_wave_barrier();

		bool stop = false;
		if (!P.approx_max) { // exact max over the 32-bit H row
			int32_t max_H, Hen0, Hst0;
			int max_t;
			if (r > 0) {
				if (packed) argmax_decode8(wave_max_u32(bh), st0, en0, en1, max_H, max_t);
				else argmax_reduce(bh, bp, st0, en0, en1, max_H, max_t);
				Hen0 = __builtin_amdgcn_readlane(h_en0, (en0 - st) & 63);
				Hst0 = __builtin_amdgcn_readlane(h_st0, (st0 - st) & 63);
			} else {
				max_H = DUAL ? (int32_t)V[0] - P.qe0 : (int32_t)(uint8_t)V[0] - P.qe - P.qe;
				max_t = 0;
				if (lane == 0) H[0] = max_H;
				Hen0 = Hst0 = max_H;
			}
			stop = finish_exact<DUAL>(P, z, r, st0, en0, max_H, max_t, Hen0, Hst0, last_max_H);
		} else {
			stop = finish_approx<DUAL>(P, z, r, st0, en0, H0, last_H0_t, [&](int t) { return V[t & CM]; },
			                           [&](int t) { return U[t & CM]; });
		}
		if (GLOBAL) __threadfence_block();
		__builtin_amdgcn_wave_barrier();
		if (stop) break;
		last_st = st, last_en = en;
	}
	__syncthreads();
	const int n_cigar = backtrack<DUAL, GLOBAL>(P, z, jb, a, lds, slab, 12 * CAP);
	write_result(a, jid, z, n_cigar, P.unplanned ? -1 : cells);
	__syncthreads();
}

// ---- the extension step's jobs ----------------------------------------------------------------
//
// Jobs with a CIGAR, the exact max, the standard nt4 scoring and a band of at most 224 cells run in
// their own kernel on a ring of kFastCap slots whose array offsets are compile-time constants.
// Instead of testing every lane for indices entering the window, the ring starts with the initial
// values (ksw2_extd2_sse.c:101-111) and the slots of the 16 indices that leave the window when st
// moves up are reset at once: the window [st-1, end] spans at most nb + 16 < kFastCap indices, so a
// slot is reset before its next index arrives.  The target and the reversed query are staged in LDS
// with zero padding (the bytes ksw2's unaligned score loads see past the sequences), a block trip
// ends at max(en, shi) instead of en + 15, and the stores are predicated instead of selected.
// run_job_lds is the reference for the semantics.

constexpr int kFastCap = 256, kFastMask = kFastCap - 1;

// a value every lane read from the same LDS address, marked uniform so that the row loop's control
// flow stays scalar
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
constexpr int kFastMaxNb = kFastCap - 32;
constexpr int kFastState = 12 * kFastCap;     // U V X Y X2 Y2 S (int8) and H (int32)

// the argmax key of the fast kernel: (H + 2^22) << 9 | rank, rank 511 for en0, 508 - (class << 6 |
// position) for body cells (class (t - st0) % 4, position (t - st0) / 4, at most 61), 2 - offset
// for the tail cells; |H| < 2^22 holds for qlen + tlen <= 30000
__device__ __forceinline__ void argmax_decode9(uint32_t k, int st0, int en0, int en1, int32_t &max_H, int &max_t)
{
	max_H = (int32_t)(k >> 9) - (1 << 22);
	const int rank = (int)(k & 511u), v = 508 - rank;
	max_t = rank == 511 ? en0 : rank >= 255 ? st0 + ((v & 63) << 2) + (v >> 6) : en1 + (2 - rank);
}

template <bool DUAL>
__device__ __forceinline__ bool fast_job(const JobP<DUAL> &P, int qt_cap)
{
	return !P.skip && P.with_cigar && !P.approx_max && !P.generic && P.nb <= kFastMaxNb && P.qlen + P.tlen <= 30000 &&
	       ((P.qlen + 15) & ~15) + ((P.tlen + 15) & ~15) <= qt_cap;
}

template <bool DUAL>
__device__ __forceinline__ void ring_reset(uint8_t *lds, int slot, int8_t ini, int8_t ini2)
{
	int8_t *R = (int8_t*)lds;
	R[slot] = ini, R[kFastCap + slot] = ini, R[2 * kFastCap + slot] = ini, R[3 * kFastCap + slot] = ini;
	if (DUAL) R[4 * kFastCap + slot] = ini2, R[5 * kFastCap + slot] = ini2;
	R[6 * kFastCap + slot] = 0;
	((int32_t*)(lds + 8 * kFastCap))[slot] = kNegInf;
}

template <bool DUAL, bool RIGHT>
__device__ __forceinline__ void run_job_fast(const DpArgs &a, int jid, const DpJob &jb, const JobP<DUAL> &P, uint8_t *lds, uint8_t *slab)
{
	const int lane = threadIdx.x;
	int8_t *U = (int8_t*)lds, *V = U + kFastCap, *X = V + kFastCap, *Y = X + kFastCap, *X2 = Y + kFastCap, *Y2 = X2 + kFastCap;
	int8_t *S = Y2 + kFastCap;
	int32_t *H = (int32_t*)(lds + 8 * kFastCap);
	const int qlen = P.qlen, tlen = P.tlen, tpad = P.tl_hi + 1;
	// scalar copies: a select between members of P would become a select of their addresses
	const int8_t m1 = P.m1, sc_N = P.sc_N, sc_mch = P.sc_mch, sc_mis = P.sc_mis;
	uint8_t *Ts = lds + kFastState, *Qr = Ts + tpad + 16;
	{
		const uint8_t *Qg = a.qseq + jb.qoff, *Tg = a.tseq + jb.toff;
		for (int i = lane; i < tpad + 16; i += 64) Ts[i] = i < tlen ? Tg[i] : 0;
		for (int i = lane; i < qlen + 16; i += 64) Qr[i] = i < qlen ? Qg[qlen - 1 - i] : 0;
		for (int i = lane; i < kFastCap; i += 64) ring_reset<DUAL>(lds, i, P.ini, P.ini2);
		__syncthreads();
	}
	DpState z;
	init_state(z);
	int cells = 0;
	int last_st = -1, last_en = -1, last_max_H = 0, prev_st = 0;
	for (int r = 0; r < qlen + tlen - 1; ++r) {
		int st0, en0;
		if (!row_range(r, qlen, tlen, P.w, st0, en0)) { z.zdropped = 1; break; }
		cells += en0 - st0 + 1;
		const int st = st0 >> 4 << 4, en = ((en0 + 16) >> 4 << 4) - 1;
		if (st != prev_st) { // indices [prev_st-1, st-1) leave the window
			for (int p = prev_st - 1 + lane; p < st - 1; p += 64)
				if (p >= 0) ring_reset<DUAL>(lds, p & kFastMask, P.ini, P.ini2);
			prev_st = st;
		}
		int8_t x1, v1, x21;
		if (st > 0 && st - 1 >= last_st && st - 1 <= last_en) {
			const int sl = (st - 1) & kFastMask;
			x1 = (int8_t)uni(X[sl]), v1 = (int8_t)uni(V[sl]), x21 = DUAL ? (int8_t)uni(X2[sl]) : 0;
		} else row_boundary<DUAL>(P, r, st, x1, v1, x21);
		const int ep = en0 > 0 ? en0 - 1 : en0;
		const int32_t Hprev = uni(H[ep & kFastMask]);  // H[en0-1] before this row's update
		if (en >= r && lane == 0) {                  // ksw2_extd2_sse.c:156-161
			int8_t yr, ur, y2r;
			row_yu<DUAL>(P, r, yr, ur, y2r);
			Y[r & kFastMask] = yr, U[r & kFastMask] = ur;
			if (DUAL) Y2[r & kFastMask] = y2r;
		}
		__builtin_amdgcn_wave_barrier();
		int shi = st0 + ((en0 - st0) >> 4 << 4) + 15;
		if (shi > P.tl_hi) shi = P.tl_hi;
		const int end = en > shi ? en : shi;         // the last index this row writes
		const int en1 = st0 + ((en0 - st0) >> 2 << 2);
		const bool exact = r > 0;
		uint32_t bk = 0;
		uint8_t *prow = slab + (size_t)r * P.nb;        // cell t at prow[t - st]
		const uint8_t *qrow = Qr + (qlen - 1 - r);   // qrow[t] = query[r - t] for t <= r, 0 up to r + 15
		int cx = x1, cv = v1, cx2 = x21;
		for (int base = st; base <= end; base += 64) {
			const int t = base + lane, sl = t & kFastMask;
			int8_t ox = X[sl], ov = V[sl], ou = U[sl], oy = Y[sl], zs = S[sl];
			int8_t ox2 = DUAL ? X2[sl] : 0, oy2 = DUAL ? Y2[sl] : 0;
			const int32_t oh = H[sl];
			const int tc = t < end ? t : end;
			const int8_t ta = (int8_t)Ts[tc], qa = (int8_t)qrow[tc];
			const int8_t fs = ((ta == m1) | (qa == m1)) ? sc_N : ta == qa ? sc_mch : sc_mis;
			zs = (unsigned)(t - st0) <= (unsigned)(shi - st0) ? fs : zs;
			int8_t xt1 = (int8_t)wave_shr1_fill((int)ox, cx), vt1 = (int8_t)wave_shr1_fill((int)ov, cv);
			int8_t x2t1 = DUAL ? (int8_t)wave_shr1_fill((int)ox2, cx2) : 0;
			if (!DUAL) {
				// _mm_cvtsi32_si128 of a negative int8 also sets bytes 1..3 (ksw2_extz2_sse.c:138-139)
				const bool qk = base == st && lane >= 1 && lane <= 3;
				xt1 = qk && x1 < 0 ? (int8_t)0xff : xt1;
				vt1 = qk && v1 < 0 ? (int8_t)0xff : vt1;
			}
			cx = __builtin_amdgcn_readlane((int)ox, 63);
			cv = __builtin_amdgcn_readlane((int)ov, 63);
			if (DUAL) cx2 = __builtin_amdgcn_readlane((int)ox2, 63);
			int8_t nU, nV, nX, nY, nX2, nY2;
			const uint8_t d = dp_cell_cg<DUAL, RIGHT>(P, xt1, vt1, x2t1, ou, oy, oy2, zs, nU, nV, nX, nY, nX2, nY2);
			const bool band = exact && t >= st0 && t <= en0;
			const int32_t hc = t < en0 ? oh + h_term<DUAL>(P, nV) : Hprev + h_term<DUAL>(P, en0 > 0 ? nU : nV);
			const int dd = t - st0;
			const uint32_t rank = t >= en0 ? 511u : t < en1 ? (uint32_t)(508 - (((dd & 3) << 6) | (dd >> 2))) : (uint32_t)(2 + en1 - t);
			const uint32_t key = ((uint32_t)(hc + (1 << 22)) << 9) | rank;
			bk = band && key > bk ? key : bk;
			// every lane stores: lanes past en write back what they loaded (the ring never maps two
			// indices of one trip to a slot), and the CIGAR bytes past en land beyond the row's
			// [st, en], which the next rows overwrite and the backtrack never reads
			const bool core = t <= en;
			S[sl] = zs;
			U[sl] = core ? nU : ou, V[sl] = core ? nV : ov, X[sl] = core ? nX : ox, Y[sl] = core ? nY : oy;
			if (DUAL) X2[sl] = core ? nX2 : ox2, Y2[sl] = core ? nY2 : oy2;
			prow[(uint32_t)(t - st)] = d;
			H[sl] = band ? hc : oh;
		}
		__builtin_amdgcn_wave_barrier();
		int32_t max_H;
		int max_t;
		if (r > 0) {
			argmax_decode9(wave_max_u32(bk), st0, en0, en1, max_H, max_t);
		} else {
			max_H = DUAL ? uni(V[0]) - P.qe0 : uni((uint8_t)V[0]) - P.qe - P.qe;
			max_t = 0;
			if (lane == 0) H[0] = max_H;
		}
		__builtin_amdgcn_wave_barrier();
		// finish_exact with H[en0] and H[st0] read only in the rows that use them
		const bool at_tend = en0 == tlen - 1;
		int32_t Hen0 = max_H;
		if (at_tend) {
			if (r > 0) Hen0 = uni(H[en0 & kFastMask]);
			if (Hen0 > z.mte) z.mte = Hen0, z.mte_q = r - en0;
		}
		if (r - st0 == qlen - 1) {
			const int32_t Hst0 = r > 0 ? uni(H[st0 & kFastMask]) : max_H;
			if (Hst0 > z.mqe) z.mqe = Hst0, z.mqe_t = st0;
		}
		if (apply_zdrop(z, max_H, r, max_t, P.zdrop, DUAL ? P.e2 : P.e)) break;
		if (at_tend && r == qlen + tlen - 2) z.score = Hen0;
		if (DUAL && (P.flag & 0x40)) { // ksw2_extd2_sse.c:361-367
			const int32_t rH = last_max_H > max_H ? last_max_H : max_H;
			const int32_t rq = qlen - (r - st0), rt = tlen - en0;
			const int32_t rm = rq >= tlen - st0 ? tlen - st0 : rt >= qlen - (r - en0) ? qlen - (r - en0) : tlen + qlen - 1 - r;
			if (rH + rm * P.max_sc + P.end_bonus < z.max) break;
		}
		last_max_H = max_H;
		last_st = st, last_en = en;
	}
	__syncthreads();
	const int n_cigar = backtrack<DUAL>(P, z, jb, a, lds, slab, kFastState + a.qt_cap + 32);
	write_result(a, jid, z, n_cigar, cells);
	__syncthreads();
}

// ---- wide bands: a team of waves per job ----------------------------------------------------------
//
// Jobs of the fast kernel's kind whose band is wider than its ring (long-read gap fills, bands of
// hundreds of cells) run on kTeamWaves waves of one workgroup.  The row state is double-buffered by
// row parity: row r reads buffer (r-1)&1 and writes buffer r&1, so the waves of a row never wait on
// each other and one barrier ends each row.  A lane reads its t-1 neighbours from the old buffer
// instead of a DPP shift.  Each row also carries the 32 indices past its end into the new buffer:
// the next row reads at most that far beyond, and no index past end + 15 has ever held computed
// values.  Every wave computes the row's finish from the same LDS values, so control flow stays
// uniform across the workgroup.  Slots of indices that leave the window are reset at the start of a
// row; they lie below st0 of the previous row, the lowest index a wave still finishing that row reads.
// run_job_lds is the reference for the semantics.

constexpr int kTeamWaves = 8, kTeamThreads = 64 * kTeamWaves;

template <bool DUAL>
__device__ __forceinline__ bool team_job(const JobP<DUAL> &P, int tcap)
{
	return tcap > 0 && !P.skip && P.with_cigar && !P.approx_max && !P.generic && P.nb > kFastMaxNb && P.nb + 144 <= tcap &&
	       P.qlen + P.tlen <= 30000;
}

template <bool DUAL>
__device__ __forceinline__ void team_reset(uint8_t *buf, int TC, int slot, int8_t ini, int8_t ini2)
{
	int8_t *R = (int8_t*)buf;
	R[slot] = ini, R[TC + slot] = ini, R[2 * TC + slot] = ini, R[3 * TC + slot] = ini;
	if (DUAL) R[4 * TC + slot] = ini2, R[5 * TC + slot] = ini2;
	R[6 * TC + slot] = 0;
	((int32_t*)(buf + 8 * TC))[slot] = kNegInf;
}

template <bool DUAL, bool RIGHT>
__device__ __forceinline__ void run_job_team(const DpArgs &a, int jid, const DpJob &jb, const JobP<DUAL> &P, uint8_t *lds, uint8_t *slab)
{
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int TC = a.tcap, TM = TC - 1;
	uint32_t *keys = (uint32_t*)(lds + 24 * TC);                // [2][kTeamWaves]
	const int qlen = P.qlen, tlen = P.tlen;
	const uint8_t *Qg = a.qseq + jb.qoff, *Tg = a.tseq + jb.toff;
	const int8_t m1 = P.m1, sc_N = P.sc_N, sc_mch = P.sc_mch, sc_mis = P.sc_mis;
	for (int i = tid; i < TC; i += kTeamThreads) {
		team_reset<DUAL>(lds, TC, i, P.ini, P.ini2);
		team_reset<DUAL>(lds + 12 * TC, TC, i, P.ini, P.ini2);
	}
	__syncthreads();
	DpState z;
	init_state(z);
	int cells = 0;
	int last_st = -1, last_en = -1, last_max_H = 0, prev_st = 0;
	for (int r = 0; r < qlen + tlen - 1; ++r) {
		// buffer offsets computed from the LDS base keep the accesses in LDS instructions
		uint8_t *bo = lds + ((r + 1) & 1) * 12 * TC, *bn = lds + (r & 1) * 12 * TC;
		const int8_t *Uo = (const int8_t*)bo, *Vo = Uo + TC, *Xo = Vo + TC, *Yo = Xo + TC, *X2o = Yo + TC, *Y2o = X2o + TC;
		const int8_t *So = Y2o + TC;
		const int32_t *Ho = (const int32_t*)(bo + 8 * TC);
		int8_t *Un = (int8_t*)bn, *Vn = Un + TC, *Xn = Vn + TC, *Yn = Xn + TC, *X2n = Yn + TC, *Y2n = X2n + TC, *Sn = Y2n + TC;
		int32_t *Hn = (int32_t*)(bn + 8 * TC);
		int st0, en0;
		if (!row_range(r, qlen, tlen, P.w, st0, en0)) { z.zdropped = 1; break; }
		cells += en0 - st0 + 1;
		const int st = st0 >> 4 << 4, en = ((en0 + 16) >> 4 << 4) - 1;
		if (st != prev_st) { // indices [prev_st-1, st-1) leave the window
			for (int p = prev_st - 1 + tid; p < st - 1; p += kTeamThreads)
				if (p >= 0) team_reset<DUAL>(lds, TC, p & TM, P.ini, P.ini2), team_reset<DUAL>(lds + 12 * TC, TC, p & TM, P.ini, P.ini2);
			prev_st = st;
		}
		int8_t x1, v1, x21;
		if (st > 0 && st - 1 >= last_st && st - 1 <= last_en) {
			const int sl = (st - 1) & TM;
			x1 = (int8_t)uni(Xo[sl]), v1 = (int8_t)uni(Vo[sl]), x21 = DUAL ? (int8_t)uni(X2o[sl]) : 0;
		} else row_boundary<DUAL>(P, r, st, x1, v1, x21);
		const int ep = en0 > 0 ? en0 - 1 : en0;
		const int32_t Hprev = uni(Ho[ep & TM]);
		int8_t yr, ur, y2r;
		row_yu<DUAL>(P, r, yr, ur, y2r);
		const int rr = en >= r ? r : -1;
		int shi = st0 + ((en0 - st0) >> 4 << 4) + 15;
		if (shi > P.tl_hi) shi = P.tl_hi;
		const int end = en > shi ? en : shi, cend = end + 32;
		const int en1 = st0 + ((en0 - st0) >> 2 << 2);
		const bool exact = r > 0;
		uint32_t bk = 0;
		uint8_t *prow = slab + (size_t)r * P.nb;
		for (int base = st + 64 * wave; base <= cend; base += 64 * kTeamWaves) {
			const int t = base + lane, sl = t & TM, sm = (t - 1) & TM;
			int8_t ox = Xo[sl], ov = Vo[sl], ou = Uo[sl], oy = Yo[sl], zs = So[sl];
			int8_t ox2 = DUAL ? X2o[sl] : 0, oy2 = DUAL ? Y2o[sl] : 0;
			const int32_t oh = Ho[sl];
			const bool first = t == st;
			int8_t xt1 = first ? x1 : Xo[sm], vt1 = first ? v1 : Vo[sm];
			int8_t x2t1 = DUAL ? (first ? x21 : X2o[sm]) : 0;
			if (!DUAL) {
				// _mm_cvtsi32_si128 of a negative int8 also sets bytes 1..3 (ksw2_extz2_sse.c:138-139)
				const bool qk = t > st && t <= st + 3;
				xt1 = qk && x1 < 0 ? (int8_t)0xff : xt1;
				vt1 = qk && v1 < 0 ? (int8_t)0xff : vt1;
			}
			oy = t == rr ? yr : oy, ou = t == rr ? ur : ou, oy2 = t == rr ? y2r : oy2;
			const int tc = t < tlen - 1 ? t : tlen - 1, qi = r - t;
			const int8_t ta0 = (int8_t)Tg[tc > 0 ? tc : 0], qa0 = (int8_t)Qg[qi < 0 ? 0 : qi < qlen ? qi : qlen - 1];
			const int8_t ta = t < tlen ? ta0 : 0, qa = qi >= 0 ? qa0 : 0;
			const int8_t fs = ((ta == m1) | (qa == m1)) ? sc_N : ta == qa ? sc_mch : sc_mis;
			zs = (unsigned)(t - st0) <= (unsigned)(shi - st0) ? fs : zs;
			int8_t nU, nV, nX, nY, nX2, nY2;
			const uint8_t d = dp_cell_cg<DUAL, RIGHT>(P, xt1, vt1, x2t1, ou, oy, oy2, zs, nU, nV, nX, nY, nX2, nY2);
			const bool band = exact && t >= st0 && t <= en0;
			const int32_t hc = t < en0 ? oh + h_term<DUAL>(P, nV) : Hprev + h_term<DUAL>(P, en0 > 0 ? nU : nV);
			const int dd = t - st0;
			const uint32_t rank = t >= en0 ? 511u : t < en1 ? (uint32_t)(508 - (((dd & 3) << 6) | (dd >> 2))) : (uint32_t)(2 + en1 - t);
			const uint32_t key = ((uint32_t)(hc + (1 << 22)) << 9) | rank;
			bk = band && key > bk ? key : bk;
			// row 0 keeps H[0] = v - (q+e) (ksw2_extd2_sse.c:339-343)
			const int32_t h0 = DUAL ? (int32_t)nV - P.qe0 : (int32_t)(uint8_t)nV - P.qe - P.qe;
			if (t <= cend) {
				const bool core = t <= en;
				Sn[sl] = zs;
				Un[sl] = core ? nU : ou, Vn[sl] = core ? nV : ov, Xn[sl] = core ? nX : ox, Yn[sl] = core ? nY : oy;
				if (DUAL) X2n[sl] = core ? nX2 : ox2, Y2n[sl] = core ? nY2 : oy2;
				Hn[sl] = band ? hc : r == 0 && t == 0 ? h0 : oh;
				prow[(uint32_t)(t - st)] = d;
			}
		}
		const uint32_t wk = wave_max_u32(bk);
		if (lane == 0) keys[(r & 1) * kTeamWaves + wave] = wk;
		__syncthreads();
		int32_t max_H;
		int max_t;
		if (r > 0) {
			const uint32_t k = wave_max_u32(lane < kTeamWaves ? keys[(r & 1) * kTeamWaves + lane] : 0u);
			argmax_decode9(k, st0, en0, en1, max_H, max_t);
		} else {
			max_H = uni(Hn[0]);
			max_t = 0;
		}
		const bool at_tend = en0 == tlen - 1;
		int32_t Hen0 = max_H;
		if (at_tend) {
			if (r > 0) Hen0 = uni(Hn[en0 & TM]);
			if (Hen0 > z.mte) z.mte = Hen0, z.mte_q = r - en0;
		}
		if (r - st0 == qlen - 1) {
			const int32_t Hst0 = r > 0 ? uni(Hn[st0 & TM]) : max_H;
			if (Hst0 > z.mqe) z.mqe = Hst0, z.mqe_t = st0;
		}
		if (apply_zdrop(z, max_H, r, max_t, P.zdrop, DUAL ? P.e2 : P.e)) break;
		if (at_tend && r == qlen + tlen - 2) z.score = Hen0;
		if (DUAL && (P.flag & 0x40)) { // ksw2_extd2_sse.c:361-367
			const int32_t rH = last_max_H > max_H ? last_max_H : max_H;
			const int32_t rq = qlen - (r - st0), rt = tlen - en0;
			const int32_t rm = rq >= tlen - st0 ? tlen - st0 : rt >= qlen - (r - en0) ? qlen - (r - en0) : tlen + qlen - 1 - r;
			if (rH + rm * P.max_sc + P.end_bonus < z.max) break;
		}
		last_max_H = max_H;
		last_st = st, last_en = en;
	}
	__syncthreads();
	const int n_cigar = backtrack<DUAL, false, kTeamThreads>(P, z, jb, a, lds, slab, 24 * TC);
	write_result(a, jid, z, n_cigar, cells);
	__syncthreads();
}

// the deferred jobs that team_job takes, one workgroup per job; the index comes through LDS
template <bool DUAL>
__global__ void __launch_bounds__(kTeamThreads) dp_team_kernel(DpArgs a)
{
	extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
	__shared__ int s_take;
	const int n = __builtin_amdgcn_readfirstlane((int)*(volatile unsigned int*)(a.next + 1));
	if (n == 0) return;
	uint8_t *slab = a.pbuf + (size_t)blockIdx.x * a.slab;
	auto take = [&]() {
		if (threadIdx.x == 0) s_take = (int)atomicAdd(a.next + 4, 1u);
		__syncthreads();
		const int v = __builtin_amdgcn_readfirstlane(s_take);
		__syncthreads();
		return v;
	};
	for (int i = take(); i < n; i = take()) {
		const int j = __builtin_amdgcn_readfirstlane(a.defer[i]);
		const DpJob jb = a.jobs[j];
		const JobP<DUAL> P = job_params<DUAL>(jb, a);
		if (team_job<DUAL>(P, a.tcap)) {
			if (P.right) run_job_team<DUAL, true>(a, j, jb, P, lds, slab);
			else run_job_team<DUAL, false>(a, j, jb, P, lds, slab);
		}
	}
}

// the next index from a shared counter, one per wave: lane 0 takes it, every lane receives it
__device__ __forceinline__ int wave_take(unsigned int *ctr)
{
	int k = 0;
	if (threadIdx.x == 0) k = (int)atomicAdd(ctr, 1u);
	return __builtin_amdgcn_readfirstlane(k);
}

// job counters in a.next: [0] taken by dp_fast_kernel, [1] jobs it deferred (listed in a.defer),
// [2], [3] and [4] deferred jobs taken by dp_kernel<DUAL, false>, dp_kernel<DUAL, true> and
// dp_team_kernel.  The job
// loops take the next index in the loop condition and have no other exit, so that the lane-0
// regions inside the body cannot become part of the loop's exit.
// With a.inline_general (the batch's widest LDS ring fits the fast kernel's LDS), the jobs the ring
// path of dp_kernel<DUAL, false> would take run here as they come, so that the batch's two job
// classes share one launch and its tail instead of running one launch after the other; only
// device-memory rings and team jobs are deferred then.
template <bool DUAL>
__global__ void __launch_bounds__(64) dp_fast_kernel(DpArgs a)
{
	extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
	__shared__ int8_t mat[32];
	const bool inl = a.inline_general != 0;
	if (inl) {
		if (threadIdx.x < 32) mat[threadIdx.x] = a.mat[threadIdx.x];
		__syncthreads();
	}
	uint8_t *slab = a.pbuf + (size_t)blockIdx.x * a.slab;
	int n = a.n_jobs;
	if (a.n_jobs_dev) { const unsigned long long nd = *a.n_jobs_dev; n = nd < (unsigned long long)n ? (int)nd : n; }
	n = __builtin_amdgcn_readfirstlane(n);
	for (int j = wave_take(a.next); j < n; j = wave_take(a.next)) {
		const DpJob jb = a.jobs[j];
		const JobP<DUAL> P = job_params<DUAL>(jb, a);
		if (!fast_job<DUAL>(P, a.qt_cap)) {
			if (inl && P.nb + 17 <= a.cap && !team_job<DUAL>(P, a.tcap)) {    // dp_kernel<DUAL, false>'s jobs
				if (((P.qlen + 15) & ~15) + ((P.tlen + 15) & ~15) <= a.qt_cap) run_job_lds<DUAL, true>(a, j, jb, P, lds, mat, slab, a.cap);
				else run_job_lds<DUAL, false>(a, j, jb, P, lds, mat, slab, a.cap);
			} else {
				const int pos = wave_take(a.next + 1);
				if (threadIdx.x == 0) a.defer[pos] = j;
			}
		} else if (P.right) run_job_fast<DUAL, true>(a, j, jb, P, lds, slab);
		else run_job_fast<DUAL, false>(a, j, jb, P, lds, slab);
	}
}

// the deferred jobs: WIDE = false takes those whose band fits the LDS ring (nb + 17 <= cap),
// WIDE = true the others, on a 12*gcap-byte ring in this block's device-memory region
template <bool DUAL, bool WIDE>
__global__ void __launch_bounds__(64) dp_kernel(DpArgs a)
{
	extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
	__shared__ int8_t mat[32];
	if (threadIdx.x < 32) mat[threadIdx.x] = a.mat[threadIdx.x];
	__syncthreads();
	uint8_t *slab = a.pbuf + (size_t)blockIdx.x * a.slab;
	// written by dp_fast_kernel; read once, as a wave-uniform value (a divergent count would turn
	// the job loop, and the barriers and lane-0 stores inside it, into divergent control flow)
	const int n = __builtin_amdgcn_readfirstlane((int)*(volatile unsigned int*)(a.next + 1));
	if (n == 0) return;
	unsigned int *next = a.next + (WIDE ? 3 : 2);
	for (int i = wave_take(next); i < n; i = wave_take(next)) {
		const int j = __builtin_amdgcn_readfirstlane(a.defer[i]);
		const DpJob jb = a.jobs[j];
		const JobP<DUAL> P = job_params<DUAL>(jb, a);
		const bool wide = P.nb + 17 > a.cap;
		if (WIDE && wide) {
			run_job_lds<DUAL, false, true>(a, j, jb, P, a.gstate + (size_t)blockIdx.x * 12 * a.gcap, mat, slab, a.gcap);
		} else if (!WIDE && !wide && !team_job<DUAL>(P, a.tcap)) {
			if (((P.qlen + 15) & ~15) + ((P.tlen + 15) & ~15) <= a.qt_cap) run_job_lds<DUAL, true>(a, j, jb, P, lds, mat, slab, a.cap);
			else run_job_lds<DUAL, false>(a, j, jb, P, lds, mat, slab, a.cap);
		}
	}
}

int dp_team_cap(int nb_max)
{
	if (nb_max <= kFastMaxNb) return 0;
	int c = 256;
	while (c < nb_max + 144 && c < 2048) c <<= 1;   // 24*2048 bytes of LDS: within 64 KB per workgroup
	return c;
}

hipError_t launch_dp(bool dual, const DpArgs &a0, unsigned grid, hipStream_t st)
{
	DpArgs a = a0;
	a.inline_general = 12 * a.cap <= kFastState;   // the ring path's state and staging fit the fast kernel's LDS
	const size_t fshm = (size_t)kFastState + (size_t)a.qt_cap + 32;
	if (dual) hipLaunchKernelGGL(dp_fast_kernel<true>, dim3(grid), dim3(64), fshm, st, a);
	else hipLaunchKernelGGL(dp_fast_kernel<false>, dim3(grid), dim3(64), fshm, st, a);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	if (a.tcap > 0) {
		const size_t tshm = (size_t)24 * a.tcap + 8 * kTeamWaves;
		// one workgroup per team job (a slab each: grid <= the slabs allocated), at most 4096
		const unsigned tgrid = grid < 4096 ? grid : 4096;
		if (dual) hipLaunchKernelGGL(dp_team_kernel<true>, dim3(tgrid), dim3(kTeamThreads), tshm, st, a);
		else hipLaunchKernelGGL(dp_team_kernel<false>, dim3(tgrid), dim3(kTeamThreads), tshm, st, a);
		e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	if (!a.inline_general) {   // otherwise the fast kernel took these jobs itself
		const size_t shm = (size_t)12 * a.cap + (size_t)a.qt_cap;
		if (dual) hipLaunchKernelGGL((dp_kernel<true, false>), dim3(grid), dim3(64), shm, st, a);
		else hipLaunchKernelGGL((dp_kernel<false, false>), dim3(grid), dim3(64), shm, st, a);
		e = hipGetLastError();
	}
	if (e != hipSuccess || a.gcap == 0) return e;
	if (dual) hipLaunchKernelGGL((dp_kernel<true, true>), dim3(grid), dim3(64), 0, st, a);
	else hipLaunchKernelGGL((dp_kernel<false, true>), dim3(grid), dim3(64), 0, st, a);
	return hipGetLastError();
}

} // namespace mbg
