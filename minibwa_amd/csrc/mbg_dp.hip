// mbg_dp.hip -- banded dual-affine (ksw_extd2) and single-affine (ksw_extz2) extension DP for gfx950.
//
// One 64-lane wavefront per job; each anti-diagonal r is computed by the lanes in parallel (cell t
// of the row on lane (t - st) % 64), the t-1 neighbour of the previous row arrives by a lane
// shuffle, and the row's state lives in LDS.  The results must equal ksw_extd2_sse
// (ksw2_extd2_sse.c:27-403) and ksw_extz2_sse (ksw2_extz2_sse.c:16-296) bit for bit, so the kernel
// keeps their observable semantics rather than a textbook recurrence:
//   * int8 difference arithmetic with wrap-around, and the unsigned/signed max mix of extz2;
//   * every row covers the 16-aligned range [st0/16*16, (en0+16)/16*16-1]: cells outside the exact
//     band are computed, kept and later read, exactly as the 16-lane SSE blocks do;
//   * the score row is written in 16-cell groups starting at st0 (so it runs past en0 with the
//     target's zero padding) and cells it does not reach keep their stale scores;
//   * the per-index state lives in a circular LDS buffer of `cap` bytes per array; an index is set
//     to the reference's initial value when it first enters the window (cap > row width + 16, so
//     no live index is overwritten);
//   * the exact-max argmax reproduces the SSE reduction order (H[en0] first, four lanes with
//     t = st0 mod 4 classes keeping their first strict maximum, merged in lane order, then the
//     scalar tail), encoded as one 64-bit key per cell and a wave max-reduction;
//   * the backtrack walks the rotated matrix (ksw2.h:140-172) from LDS-staged chunks of rows.
#include "mbg_internal.h"

namespace mbg {

constexpr int32_t kNegInf = -0x40000000;

// Cross-lane steps through DPP (VALU data-parallel primitives) rather than LDS permutes: the row
// recurrence is latency-bound, and a DPP move costs a few cycles where ds_bpermute costs an LDS
// round trip.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
	return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

// value of lane-1 (lane 0 receives 0)
__device__ __forceinline__ int wave_shr1(int v) { return (int)dpp_u32<0x138>((uint32_t)v); }

// unsigned max over the 64 lanes, returned as a wave-uniform value (rows of 16 first, then the
// row broadcasts fold rows 0-1 and 2-3, then lanes 31 -> 63)
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
	uint32_t m;
	m = dpp_u32<0x111>(v); v = m > v ? m : v;   // row_shr:1
	m = dpp_u32<0x112>(v); v = m > v ? m : v;   // row_shr:2
	m = dpp_u32<0x114>(v); v = m > v ? m : v;   // row_shr:4
	m = dpp_u32<0x118>(v); v = m > v ? m : v;   // row_shr:8
	m = dpp_u32<0x142, 0xa>(v); v = m > v ? m : v;   // row_bcast:15
	m = dpp_u32<0x143, 0xc>(v); v = m > v ? m : v;   // row_bcast:31
	return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

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

template <bool DUAL>
__device__ void run_job(const DpArgs &a, int jid, uint8_t *lds, const int8_t *mat, uint8_t *slab)
{
	const int lane = threadIdx.x;
	const int CAP = a.cap, CM = CAP - 1;
	int8_t *U = (int8_t*)lds, *V = U + CAP, *X = V + CAP, *Y = X + CAP, *X2 = Y + CAP, *Y2 = X2 + CAP, *S = Y2 + CAP;
	int32_t *H = (int32_t*)(lds + 8 * CAP);
	const DpJob jb = a.jobs[jid];
	const uint8_t *Q = a.qseq + jb.qoff, *T = a.tseq + jb.toff;
	const int qlen = jb.qlen, tlen = jb.tlen, m = a.m, flag = jb.flag, zdrop = jb.zdrop, end_bonus = jb.end_bonus;
	int8_t q = jb.q, e = jb.e, q2 = jb.q2, e2 = jb.e2;
	DpState z;
	z.max_q = z.max_t = z.mqe_t = z.mte_q = -1;
	z.max = 0, z.score = z.mqe = z.mte = kNegInf;
	z.zdropped = 0, z.reach_end = 0;
	int cells = 0, n_cigar = 0;
	const bool with_cigar = !(flag & 0x01), approx_max = !!(flag & 0x08), right = !!(flag & 0x02), generic = !!(flag & 0x04);
	bool skip = DUAL ? (m <= 1 || qlen <= 0 || tlen <= 0) : (m <= 0 || qlen <= 0 || tlen <= 0);

	// parameters (ksw2_extd2_sse.c:60-97 / ksw2_extz2_sse.c:41-74)
	const int qe0 = q + e;                       // extd2's scalar q+e is taken before the swap
	if (DUAL && q2 + e2 < q + e) { int8_t t = q; q = q2; q2 = t; t = e; e = e2; e2 = t; }
	const int8_t qe8 = (int8_t)(q + e), qe28 = (int8_t)(q2 + e2);
	const int qe = q + e;                        // extz2's scalar q+e
	const int8_t sc_mch = mat[0], sc_mis = mat[1], m1 = (int8_t)(m - 1);
	int8_t sc_N = 0, qe2x = 0, max_sc_ = 0;
	int w = jb.w, tlen_ = (tlen + 15) / 16, n_col_ = 1, max_sc = 0, min_sc = 0, long_thres = 0, long_diff = 0;
	if (!skip) {
		sc_N = DUAL ? (mat[m * m - 1] == 0 ? (int8_t)-e2 : mat[m * m - 1]) : (mat[m * m - 1] == 0 ? (int8_t)-e : mat[m * m - 1]);
		qe2x = (int8_t)((q + e) * 2);
		max_sc_ = (int8_t)(mat[0] + (q + e) * 2);
		dp_geom(qlen, tlen, w, &w, &n_col_);
		max_sc = mat[0], min_sc = mat[1];
		for (int t = 1; t < m * m; ++t) {
			max_sc = max_sc > mat[t] ? max_sc : mat[t];
			min_sc = min_sc < mat[t] ? min_sc : mat[t];
		}
		if (-min_sc > 2 * (q + e)) skip = true;
		if (DUAL) {
			long_thres = e != e2 ? (q2 - q) / (e - e2) - 1 : 0;
			if (q2 + e2 + long_thres * e2 > q + e + long_thres * e) ++long_thres;
			long_diff = long_thres * (e - e2) - (q2 - q) - e2;
		}
	}
	const int nb = n_col_ * 16;
	const int8_t ini = DUAL ? (int8_t)(-q - e) : 0, ini2 = DUAL ? (int8_t)(-q2 - e2) : 0;
	const int tl_hi = tlen_ * 16 - 1;

	int hw = -1, last_st = -1, last_en = -1, last_max_H = 0, last_H0_t = 0;
	int32_t H0 = 0;
	for (int r = 0; !skip && r < qlen + tlen - 1; ++r) {
		int st = 0, en = tlen - 1;
		if (st < r - qlen + 1) st = r - qlen + 1;
		if (en > r) en = r;
		if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
		if (en > (r + w) >> 1) en = (r + w) >> 1;
		if (st > en) { z.zdropped = 1; break; }
		const int st0 = st, en0 = en;
		cells += en0 - st0 + 1;
		st = st >> 4 << 4;
		en = ((en + 16) >> 4 << 4) - 1;
		// indices entering the window take the initial values (memset/calloc/NEG_INF)
		int hi = en + 15;
		if (hi > tl_hi) hi = tl_hi;
		if (hi > hw) {
			for (int idx = hw + 1 + lane; idx <= hi; idx += 64) {
				const int sl = idx & CM;
				U[sl] = ini; V[sl] = ini; X[sl] = ini; Y[sl] = ini;
				if (DUAL) { X2[sl] = ini2; Y2[sl] = ini2; }
				S[sl] = 0;
				H[sl] = kNegInf;
			}
			hw = hi;
			__syncthreads();
		}
		// boundary values of cell st-1 (ksw2_extd2_sse.c:141-155 / ksw2_extz2_sse.c:110-115)
		int8_t x1, x21 = 0, v1;
		if (st > 0) {
			if (st - 1 >= last_st && st - 1 <= last_en) {
				const int sl = (st - 1) & CM;
				x1 = X[sl], v1 = V[sl];
				if (DUAL) x21 = X2[sl];
			} else if (DUAL) {
				x1 = (int8_t)(-q - e), x21 = (int8_t)(-q2 - e2), v1 = (int8_t)(-q - e);
			} else x1 = v1 = 0;
		} else if (DUAL) {
			x1 = (int8_t)(-q - e), x21 = (int8_t)(-q2 - e2);
			v1 = (int8_t)(r == 0 ? -q - e : r < long_thres ? -e : r == long_thres ? long_diff : -e2);
		} else {
			x1 = 0, v1 = r ? q : 0;
		}
		if (en >= r && lane == 0) {
			const int sl = r & CM;
			if (DUAL) {
				Y[sl] = (int8_t)(-q - e), Y2[sl] = (int8_t)(-q2 - e2);
				U[sl] = (int8_t)(r == 0 ? -q - e : r < long_thres ? -e : r == long_thres ? long_diff : -e2);
			} else {
				Y[sl] = 0, U[sl] = (int8_t)(r ? q : 0);
			}
		}
		// scores (loop fission in the reference): 16-cell groups from st0, padding past tlen reads 0
		if (!generic) {
			int shi = st0 + ((en0 - st0) >> 4 << 4) + 15;
			if (shi > tl_hi) shi = tl_hi;
			for (int t = st0 + lane; t <= shi; t += 64) {
				const int8_t ta = t < tlen ? (int8_t)T[t] : 0, qa = t <= r ? (int8_t)Q[r - t] : 0;
				S[t & CM] = (ta == m1 || qa == m1) ? sc_N : ta == qa ? sc_mch : sc_mis;
			}
		} else {
			for (int t = st0 + lane; t <= en0; t += 64)
				S[t & CM] = mat[T[t] * m + Q[r - t]];
		}
		__syncthreads();

		// core: every cell of [st, en]
		const int32_t Hprev = en0 > 0 ? H[(en0 - 1) & CM] : H[en0 & CM];
		uint8_t *prow = slab + (size_t)r * nb;
		int8_t cx = x1, cv = v1, cx2 = x21;
		for (int base = st; base <= en; base += 64) {
			const int t = base + lane, sl = t & CM;
			const bool valid = t <= en;
			int8_t ox = 0, ov = 0, ox2 = 0, ou = 0, oy = 0, oy2 = 0, zs = 0;
			if (valid) {
				ox = X[sl]; ov = V[sl]; ou = U[sl]; oy = Y[sl]; zs = S[sl];
				if (DUAL) { ox2 = X2[sl]; oy2 = Y2[sl]; }
			}
			int8_t xt1 = (int8_t)wave_shr1((int)ox), vt1 = (int8_t)wave_shr1((int)ov);
			int8_t x2t1 = DUAL ? (int8_t)wave_shr1((int)ox2) : 0;
			if (lane == 0) { xt1 = cx; vt1 = cv; x2t1 = cx2; }
			if (!DUAL && base == st && lane >= 1 && lane <= 3) {
				// _mm_cvtsi32_si128 of a negative int8 also sets bytes 1..3 (ksw2_extz2_sse.c:138-139)
				if (x1 < 0) xt1 = (int8_t)0xff;
				if (v1 < 0) vt1 = (int8_t)0xff;
			}
			cx = (int8_t)__builtin_amdgcn_readlane((int)ox, 63); cv = (int8_t)__builtin_amdgcn_readlane((int)ov, 63);
			if (DUAL) cx2 = (int8_t)__builtin_amdgcn_readlane((int)ox2, 63);
			if (!valid) continue;
			uint8_t d = 0;
			if (DUAL) {
				int8_t zz = zs;
				int8_t A = (int8_t)(xt1 + vt1), B = (int8_t)(oy + ou), A2 = (int8_t)(x2t1 + vt1), B2 = (int8_t)(oy2 + ou);
				if (!with_cigar) {
					zz = smax8(zz, A); zz = smax8(zz, B); zz = smax8(zz, A2); zz = smax8(zz, B2);
				} else if (!right) {
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
				zz = smin8(zz, sc_mch);
				U[sl] = (int8_t)(zz - vt1);
				V[sl] = (int8_t)(zz - ou);
				int8_t tmp = (int8_t)(zz - q);
				A = (int8_t)(A - tmp); B = (int8_t)(B - tmp);
				tmp = (int8_t)(zz - q2);
				A2 = (int8_t)(A2 - tmp); B2 = (int8_t)(B2 - tmp);
				if (!with_cigar || !right) {
					X[sl]  = (int8_t)((A > 0 ? A : 0) - qe8);
					Y[sl]  = (int8_t)((B > 0 ? B : 0) - qe8);
					X2[sl] = (int8_t)((A2 > 0 ? A2 : 0) - qe28);
					Y2[sl] = (int8_t)((B2 > 0 ? B2 : 0) - qe28);
					d |= (A > 0 ? 0x08 : 0) | (B > 0 ? 0x10 : 0) | (A2 > 0 ? 0x20 : 0) | (B2 > 0 ? 0x40 : 0);
				} else {
					X[sl]  = (int8_t)((A >= 0 ? A : 0) - qe8);
					Y[sl]  = (int8_t)((B >= 0 ? B : 0) - qe8);
					X2[sl] = (int8_t)((A2 >= 0 ? A2 : 0) - qe28);
					Y2[sl] = (int8_t)((B2 >= 0 ? B2 : 0) - qe28);
					d |= (A >= 0 ? 0x08 : 0) | (B >= 0 ? 0x10 : 0) | (A2 >= 0 ? 0x20 : 0) | (B2 >= 0 ? 0x40 : 0);
				}
			} else {
				const uint8_t uvt1 = (uint8_t)vt1, uut = (uint8_t)ou;
				int8_t zz = (int8_t)((uint8_t)zs + (uint8_t)qe2x);
				int8_t A = (int8_t)((uint8_t)xt1 + uvt1), B = (int8_t)((uint8_t)oy + uut);
				if (!with_cigar) {
					zz = smax8(zz, A);
				} else if (!right) {
					d = A > zz ? 1 : 0;
					zz = smax8(zz, A);
					d = B > zz ? 2 : d;
				} else {
					d = zz > A ? 0 : 1;
					zz = smax8(zz, A);
					d = zz > B ? d : 2;
				}
				uint8_t zu = (uint8_t)zz > (uint8_t)B ? (uint8_t)zz : (uint8_t)B;
				zu = zu < (uint8_t)max_sc_ ? zu : (uint8_t)max_sc_;
				U[sl] = (int8_t)(uint8_t)(zu - uvt1);
				V[sl] = (int8_t)(uint8_t)(zu - uut);
				const int8_t zq = (int8_t)(uint8_t)(zu - (uint8_t)q);
				A = (int8_t)(A - zq); B = (int8_t)(B - zq);
				if (!with_cigar || !right) {
					X[sl] = A > 0 ? A : 0; Y[sl] = B > 0 ? B : 0;
					d |= (A > 0 ? 0x08 : 0) | (B > 0 ? 0x10 : 0);
				} else {
					X[sl] = A >= 0 ? A : 0; Y[sl] = B >= 0 ? B : 0;
					d |= (A >= 0 ? 0x08 : 0) | (B >= 0 ? 0x10 : 0);
				}
			}
			if (with_cigar) prow[t - st] = d;
		}
		__syncthreads();

		bool stop = false;
		if (!approx_max) { // exact max over the 32-bit H row
			int32_t max_H, max_t, Hen0, Hst0;
			if (r > 0) {
				const int en1 = st0 + ((en0 - st0) >> 2 << 2);
				uint32_t bh = 0, bp = 0;         // this lane's best (H ^ sign bit, priority)
				int32_t h_first = 0;
				for (int t = st0 + lane; t < en0; t += 64) {
					const int sl = t & CM;
					const int32_t h = H[sl] + (DUAL ? (int32_t)V[sl] : (int32_t)(uint8_t)V[sl] - qe);
					H[sl] = h;
					if (t == st0) h_first = h;
					const uint32_t pri = t < en1
						? (0x80000000u | ((uint32_t)(3 - ((t - st0) & 3)) << 28) | (0x0FFFFFFFu - (uint32_t)((t - st0) >> 2)))
						: (0x7FFFFFFFu - (uint32_t)(t - en1));
					const uint32_t hk = (uint32_t)h ^ 0x80000000u;
					if (hk > bh || (hk == bh && pri > bp)) bh = hk, bp = pri;
				}
				int32_t h_en0 = 0;
				if (lane == 0) {
					const int sl = en0 & CM;
					if (DUAL) h_en0 = en0 > 0 ? Hprev + (int32_t)U[sl] : Hprev + (int32_t)V[sl];
					else h_en0 = en0 > 0 ? Hprev + (int32_t)(uint8_t)U[sl] - qe : Hprev + (int32_t)(uint8_t)V[sl] - qe;
					H[sl] = h_en0;
					const uint32_t hk = (uint32_t)h_en0 ^ 0x80000000u;
					if (hk > bh || (hk == bh && 0xFFFFFFFFu > bp)) bh = hk, bp = 0xFFFFFFFFu;
				}
				const uint32_t mh = wave_max_u32(bh);
				const uint32_t pri = wave_max_u32(bh == mh ? bp : 0u);
				max_H = (int32_t)(mh ^ 0x80000000u);
				if (pri == 0xFFFFFFFFu) max_t = en0;
				else if (pri & 0x80000000u) max_t = st0 + (int)(0x0FFFFFFFu - (pri & 0x0FFFFFFFu)) * 4 + (3 - (int)((pri >> 28) & 3));
				else max_t = en1 + (int)(0x7FFFFFFFu - pri);
				Hen0 = __builtin_amdgcn_readlane(h_en0, 0);
				Hst0 = en0 > st0 ? __builtin_amdgcn_readlane(h_first, 0) : Hen0;
			} else {
				max_H = DUAL ? (int32_t)V[0] - qe0 : (int32_t)(uint8_t)V[0] - qe - qe;
				max_t = 0;
				if (lane == 0) H[0] = max_H;
				Hen0 = Hst0 = max_H;
			}
			if (en0 == tlen - 1 && Hen0 > z.mte) z.mte = Hen0, z.mte_q = r - en0;
			if (r - st0 == qlen - 1 && Hst0 > z.mqe) z.mqe = Hst0, z.mqe_t = st0;
			if (apply_zdrop(z, max_H, r, max_t, zdrop, DUAL ? e2 : e)) stop = true;
			else {
				if (r == qlen + tlen - 2 && en0 == tlen - 1) z.score = Hen0;
				if (DUAL && (flag & 0x40)) { // ksw2_extd2_sse.c:361-367
					const int32_t rH = last_max_H > max_H ? last_max_H : max_H;
					const int32_t rq = qlen - (r - st0), rt = tlen - en0;
					const int32_t rm = rq >= tlen - st0 ? tlen - st0 : rt >= qlen - (r - en0) ? qlen - (r - en0) : tlen + qlen - 1 - r;
					if (rH + rm * max_sc + end_bonus < z.max) stop = true;
				}
				last_max_H = max_H;
			}
		} else { // approximate max (ksw2_extd2_sse.c:369-385 / ksw2_extz2_sse.c:262-278)
			if (r > 0) {
				const bool in0 = last_H0_t >= st0 && last_H0_t <= en0, in1 = last_H0_t + 1 >= st0 && last_H0_t + 1 <= en0;
				if (in0 && in1) {
					const int32_t d0 = DUAL ? (int32_t)V[last_H0_t & CM] : (int32_t)(uint8_t)V[last_H0_t & CM] - qe;
					const int32_t d1 = DUAL ? (int32_t)U[(last_H0_t + 1) & CM] : (int32_t)(uint8_t)U[(last_H0_t + 1) & CM] - qe;
					if (d0 > d1) H0 += d0;
					else H0 += d1, ++last_H0_t;
				} else if (in0) {
					H0 += DUAL ? (int32_t)V[last_H0_t & CM] : (int32_t)(uint8_t)V[last_H0_t & CM] - qe;
				} else {
					++last_H0_t;
					H0 += DUAL ? (int32_t)U[last_H0_t & CM] : (int32_t)(uint8_t)U[last_H0_t & CM] - qe;
				}
				if (!DUAL && (flag & 0x10) && apply_zdrop(z, H0, r, last_H0_t, zdrop, e)) stop = true;
			} else {
				H0 = DUAL ? (int32_t)V[0] - qe0 : (int32_t)(uint8_t)V[0] - qe - qe;
				last_H0_t = 0;
			}
			if (!stop && DUAL && (flag & 0x10) && apply_zdrop(z, H0, r, last_H0_t, zdrop, e2)) stop = true;
			if (!stop && r == qlen + tlen - 2 && en0 == tlen - 1) z.score = H0;
		}
		__syncthreads();
		if (stop) break;
		last_st = st, last_en = en;
	}

	// backtrack (ksw2_extd2_sse.c:391-402, ksw2.h:140-172)
	uint32_t *cg = a.cigar + jb.cigar_off;
	int i0 = -1, j0 = -1;
	if (!skip && with_cigar) {
		if (!z.zdropped && !(flag & 0x40)) i0 = tlen - 1, j0 = qlen - 1;
		else if ((flag & 0x40) && z.mqe + end_bonus > z.max) z.reach_end = 1, i0 = z.mqe_t, j0 = qlen - 1;
		else if (z.max_t >= 0 && z.max_q >= 0) i0 = z.max_t, j0 = z.max_q;
	}
	if (i0 >= 0 || j0 >= 0) {
		uint8_t *stg = lds;
		const int K = (12 * CAP) / nb;
		int chunk_lo = 0x7fffffff;
		int i = i0, j = j0, state = 0, last_op = -1;
		uint32_t cur = 0;
		while (i >= 0 && j >= 0) {
			const int r = i + j;
			int s0 = 0, e0 = tlen - 1;
			if (s0 < r - qlen + 1) s0 = r - qlen + 1;
			if (e0 > r) e0 = r;
			if (s0 < (r - w + 1) >> 1) s0 = (r - w + 1) >> 1;
			if (e0 > (r + w) >> 1) e0 = (r + w) >> 1;
			const int off = s0 >> 4 << 4, off_end = ((e0 + 16) >> 4 << 4) - 1;
			int force = -1;
			if (i < off) force = 2;
			if (i > off_end) force = 1;
			uint32_t tmp = 0;
			if (force < 0) {
				if (r < chunk_lo) {
					const int lo = r - K + 1 > 0 ? r - K + 1 : 0;
					const int bytes = (r - lo + 1) * nb;
					__syncthreads();
					for (int o = lane * 16; o < bytes; o += 64 * 16)
						*(uint4*)(stg + o) = *(const uint4*)(slab + (size_t)lo * nb + o);
					__syncthreads();
					chunk_lo = lo;
				}
				tmp = stg[(r - chunk_lo) * nb + i - off];
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
		if (!(flag & 0x80)) {
			__syncthreads();
			for (int k = lane; k < n_cigar / 2; k += 64) {
				const uint32_t t0 = cg[k], t1 = cg[n_cigar - 1 - k];
				cg[k] = t1, cg[n_cigar - 1 - k] = t0;
			}
		}
	}
	if (lane == 0) {
		DpRes o;
		o.max = z.max, o.zdropped = z.zdropped, o.max_q = z.max_q, o.max_t = z.max_t;
		o.mqe = z.mqe, o.mqe_t = z.mqe_t, o.mte = z.mte, o.mte_q = z.mte_q;
		o.score = z.score, o.reach_end = z.reach_end, o.n_cigar = n_cigar, o.cells = cells;
		a.res[jid] = o;
	}
	__syncthreads();
}

template <bool DUAL>
__global__ void __launch_bounds__(64) dp_kernel(DpArgs a)
{
	extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
	__shared__ int8_t mat[32];
	if (threadIdx.x < 32) mat[threadIdx.x] = a.mat[threadIdx.x];
	__syncthreads();
	uint8_t *slab = a.pbuf + (size_t)blockIdx.x * a.slab;
	int n = a.n_jobs;
	if (a.n_jobs_dev) { const unsigned long long nd = *a.n_jobs_dev; n = nd < (unsigned long long)n ? (int)nd : n; }
	for (int j = blockIdx.x; j < n; j += gridDim.x)
		run_job<DUAL>(a, j, lds, mat, slab);
}

hipError_t launch_dp(bool dual, const DpArgs &a, unsigned grid, hipStream_t st)
{
	const size_t shm = (size_t)12 * a.cap;
	if (dual) hipLaunchKernelGGL(dp_kernel<true>, dim3(grid), dim3(64), shm, st, a);
	else hipLaunchKernelGGL(dp_kernel<false>, dim3(grid), dim3(64), shm, st, a);
	return hipGetLastError();
}

} // namespace mbg
