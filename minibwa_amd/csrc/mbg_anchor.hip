// mbg_anchor.hip -- mb_anchor (seed.c:109-339, no methylation conversion) and mb_lchain_dp
// (lchain.c:25-240) on gfx950: a read's SMEM intervals become contig-local anchors, and its anchors
// chains, in the reference's order.
//
//   anc_rows_kernel      one thread per read: mb_seed_sort_dedup, then the rows to locate for each
//                        group of intervals sharing (x[0], size): all of a group of size <= max_occ,
//                        else max_occ evenly stepped ones (seed.c:304-315)
//   sa_kernel            (mbg_fm.hip) the rows' SA positions, in place
//   anc_build_kernel     one thread per read: each position of each interval of its group mapped to a
//                        contig and strand (seed.c:176-210), the anchors sorted by concatenated tpos,
//                        made contig-local, deduplicated (seed.c:140-166) and compacted
//   lchain_kernel        one thread per read: the chaining DP with its skip heuristics, the backtrack
//                        and the compaction of chained anchors (lchain.c:145-240)
//
// Every sort is klib's radix sort (ksort.h:108-161): insertion sort up to 64 elements, else an
// in-place most-significant-byte-first radix sort whose cycle permutation is not stable, so anchors
// with equal keys keep exactly the order that procedure leaves; it is restated here byte by byte in
// place of recursion (rs_sort below).
#include "mbg_internal.h"

namespace mbg {

template <class T, class K>
__device__ void rs_insert(T *a, int64_t beg, int64_t end, K key)
{
	for (int64_t i = beg + 1; i < end; ++i)
		if (key(a[i]) < key(a[i - 1])) {
			const T tmp = a[i];
			int64_t j = i;
			for (; j > beg && key(tmp) < key(a[j - 1]); --j) a[j] = a[j - 1];
			a[j] = tmp;
		}
}

// one most-significant-byte-first pass over [beg, end) on byte s of the key (ksort.h:117-148), then
// insertion sort of its buckets of 2..64 elements; bb and be hold 256 bucket bounds each
template <class T, class K>
__device__ void rs_pass(T *a, int64_t beg, int64_t end, int s, K key, int32_t *bb, int32_t *be)
{
	for (int k = 0; k < 256; ++k) bb[k] = be[k] = (int32_t)beg;
	for (int64_t i = beg; i < end; ++i) ++be[(int)(key(a[i]) >> s & 255)];
	for (int k = 1; k < 256; ++k) be[k] += be[k - 1] - beg, bb[k] = be[k - 1];
	for (int k = 0; k < 256;) {                     // the in-place cycle permutation
		if (bb[k] != be[k]) {
			int l = (int)(key(a[bb[k]]) >> s & 255);
			if (l != k) {
				T tmp = a[bb[k]];
				do {
					const T sw = tmp;
					tmp = a[bb[l]];
					a[bb[l]++] = sw;
					l = (int)(key(tmp) >> s & 255);
				} while (l != k);
				a[bb[k]++] = tmp;
			} else ++bb[k];
		} else ++k;
	}
	if (s == 0) return;                             // buckets of equal keys
	for (int k = 0; k < 256; ++k) {
		const int64_t b0 = k ? be[k - 1] : beg, m = be[k] - b0;
		if (m > 1 && m <= 64) rs_insert(a, b0, be[k], key);
	}
}

// klib's radix sort of a[0..n).  The reference recurses into every bucket of more than 64 elements
// with the next byte; those buckets are exactly the maximal runs of more than 64 elements sharing
// the key bits above the next byte (a smaller bucket was insertion-sorted whole, so no such run lies
// in it), and buckets are disjoint, so each byte is done as one sweep over such runs.
template <class T, class K>
__device__ void rs_sort(T *a, int64_t n, K key, int32_t *bb, int32_t *be)
{
	if (n <= 64) { rs_insert(a, 0, n, key); return; }
	rs_pass(a, 0, n, 56, key, bb, be);
	for (int s = 48; s >= 0; s -= 8) {
		const int hs = s + 8;
		for (int64_t i = 0; i < n;) {
			const uint64_t h = key(a[i]) >> hs;
			int64_t j = i + 1;
			while (j < n && key(a[j]) >> hs == h) ++j;
			if (j - i > 64) rs_pass(a, i, j, s, key, bb, be);
			i = j;
		}
	}
}

struct KeyX0 { __device__ uint64_t operator()(const Sai &v) const { return v.x0; } };
struct KeySize { __device__ uint64_t operator()(const Sai &v) const { return v.size; } };
struct KeyInfo { __device__ uint64_t operator()(const Sai &v) const { return v.info; } };
struct KeyTpos { __device__ uint64_t operator()(const AncRec &v) const { return (uint64_t)v.tpos; } };

// number of SA rows located for an interval group of `size` rows (seed.c:304-315)
__device__ __forceinline__ int64_t n_located(uint64_t size, int32_t max_occ)
{
	return size < (uint64_t)max_occ ? (int64_t)size : max_occ;
}

__global__ void __launch_bounds__(64) anc_rows_kernel(AncArgs a)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= a.n_reads) return;
	int32_t bb[256], be[256];
	Sai *u = a.u + a.uoff[r];
	int64_t n = a.uoff[r + 1] - a.uoff[r];
	// mb_seed_sort_dedup (seed.c:109-138): by x[0]; a run of equal x[0] by size, reversed, and its runs
	// of equal size by info; then one of each (x[0], size, info)
	if (n > 1) {
		rs_sort(u, n, KeyX0(), bb, be);
		for (int64_t i = 1, i0 = 0; i <= n; ++i) {
			if (i < n && u[i].x0 == u[i0].x0) continue;
			if (i - i0 > 1) {
				rs_sort(u + i0, i - i0, KeySize(), bb, be);
				for (int64_t k = 0; k < (i - i0) / 2; ++k) {
					const Sai t = u[i0 + k];
					u[i0 + k] = u[i - 1 - k], u[i - 1 - k] = t;
				}
				for (int64_t k = i0 + 1, k0 = i0; k <= i; ++k)
					if (k == i || u[k0].size != u[k].size) {
						if (k - k0 > 1) rs_sort(u + k0, k - k0, KeyInfo(), bb, be);
						k0 = k;
					}
			}
			i0 = i;
		}
		int64_t j = 0;
		for (int64_t i = 1; i < n; ++i)
			if (!(u[i].x0 == u[j].x0 && u[i].size == u[j].size && u[i].info == u[j].info)) u[++j] = u[i];
		n = j + 1;
	}
	a.nu[r] = n;
	// the rows to locate, group by group of equal (x[0], size): all rows of a group of at most max_occ,
	// else max_occ rows at steps of (size - j) / (max_occ - cnt), at least 1 (seed.c:304-315)
	uint64_t *rows = a.rows + a.aoff[r];
	const int64_t cap = a.aoff[r + 1] - a.aoff[r];
	int64_t m = 0;
	for (int64_t i = 1, i0 = 0; i <= n; ++i) {
		if (i < n && u[i].x0 == u[i0].x0 && u[i].size == u[i0].size) continue;
		const uint64_t size = u[i0].size;
		const int64_t ns = n_located(size, a.max_occ);
		if (m + ns > cap) {
			// every slot still gets a row: a call that reuses the workspace holds located values there
			// (row 0 locates to 2^64-1, the '$' suffix), and the SA walk must never take one for a row
			for (int64_t k = 0; k < cap; ++k) rows[k] = 0;
			a.nu[r] = -1;
			return;
		}
		if ((int64_t)size <= a.max_occ)
			for (int64_t j = 0; j < ns; ++j) rows[m++] = u[i0].x0 + (uint64_t)j;
		else
			for (int64_t j = 0, cnt = 0; j < (int64_t)size && cnt < a.max_occ; ++cnt) {
				rows[m++] = u[i0].x0 + (uint64_t)j;
				const int32_t step = (int32_t)((size - (uint64_t)j) / (uint64_t)(a.max_occ - cnt));   // an int32 there
				j += step < 1 ? 1 : step;
			}
		i0 = i;
	}
	for (int64_t k = m; k < cap; ++k) rows[k] = 0;        // row 0 is sampled: located without a walk
}

__global__ void __launch_bounds__(64) anc_build_kernel(AncArgs a)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= a.n_reads) return;
	const int64_t n = a.nu[r];
	if (n < 0) { a.nanc[r] = -1; return; }
	int32_t bb[256], be[256];
	const Sai *u = a.u + a.uoff[r];
	const uint64_t *pos = a.rows + a.aoff[r];
	AncRec *v = a.anc + a.aoff[r];
	const int64_t cap = a.aoff[r + 1] - a.aoff[r];
	const int32_t qlen = a.qlen[r];
	const uint64_t tot = a.lb.tot_len;
	const uint64_t *coff = a.lb.ctg_off, *clen = a.lb.ctg_len;
	int64_t na = 0, m = 0;
	for (int64_t i = 1, i0 = 0; i <= n; ++i) {
		if (i < n && u[i].x0 == u[i0].x0 && u[i].size == u[i0].size) continue;
		const int64_t ns = n_located(u[i0].size, a.max_occ);
		for (int64_t k = 0; k < ns; ++k) {
			const uint64_t p = pos[m++];
			for (int64_t j = i0; j < i; ++j) {
				const int32_t qs = (int32_t)(u[j].info >> 32), qe = (int32_t)u[j].info, len = qe - qs;
				const uint64_t en = p + (uint64_t)len;
				// l2b_intv2cid / l2b_pos2cid (l2bit.c:9-34): [p, en) on one strand of one contig
				if (en > tot * 2 || (p < tot && tot < en)) continue;
				const int rev = p >= tot;
				const int64_t s = rev ? (int64_t)(tot * 2 - en) : (int64_t)p;
				int64_t lo = 0, hi = (int64_t)a.lb.n_ctg, tid = -1, cst = 0;
				while (lo < hi) {
					const int64_t mid = (lo + hi) / 2, off = (int64_t)coff[mid], cl = (int64_t)clen[mid];
					if (off <= s && s < off + cl) {
						cst = s - off;
						tid = s + len <= off + cl ? mid : -1;
						break;
					} else if (s < off) hi = mid;
					else lo = mid + 1;
				}
				if (tid < 0) continue;
				if (na == cap) { a.nanc[r] = -1; return; }
				AncRec q;
				q.sid = (int32_t)(tid << 1 | rev), q.len = len, q.flag = 0;
				q.qpos = rev ? qlen - 1 - qs : qs + len - 1;
				q.tpos = (int64_t)(coff[tid] * 2 + clen[tid] * (uint64_t)rev) + cst + len - 1;   // concatenated, for the sort
				v[na++] = q;
			}
		}
		i0 = i;
	}
	rs_sort(v, na, KeyTpos(), bb, be);
	for (int64_t i = 0; i < na; ++i) {
		const int64_t t = v[i].sid >> 1;
		v[i].tpos -= (int64_t)(coff[t] * 2 + clen[t] * (uint64_t)(v[i].sid & 1));
	}
	// mb_anchor_dedup (seed.c:140-166): within 100 predecessors on the same strand and contig, an
	// anchor starting inside another on the same diagonal is dropped, and of two ending at the same
	// (tpos, qpos) the later; the mark is the flag word's top bit (mb_anchor_t::flt)
	for (int64_t i = 1; i < na; ++i) {
		AncRec &ai = v[i];
		const int64_t tsi = ai.tpos + 1 - ai.len;
		const int32_t qsi = ai.qpos + 1 - ai.len;
		for (int64_t j = i - 1, k = 0; j >= 0 && k < 100; --j, ++k) {
			AncRec &aj = v[j];
			if (aj.sid != ai.sid || aj.tpos < tsi) break;
			const int64_t tsj = aj.tpos + 1 - aj.len;
			const int32_t qsj = aj.qpos + 1 - aj.len;
			if (tsj >= tsi) {
				if (tsj - tsi == qsj - qsi) aj.flag |= 0x80000000u;
			} else if (ai.tpos == aj.tpos) {
				if (ai.qpos == aj.qpos) ai.flag |= 0x80000000u;
			}
		}
	}
	int64_t j = 0;
	for (int64_t i = 0; i < na; ++i)
		if (!(v[i].flag >> 31)) v[j++] = v[i];
	a.nanc[r] = j;
}

// ---- chaining: mb_lchain_dp (lchain.c:145-240), one thread per read ------------------------------

struct KeyU128 { __device__ uint64_t operator()(const U128 &v) const { return v.x; } };

// comput_sc (lchain.c:125-143): the score of chaining aj -> ai, or INT32_MIN; the gap penalty is
// float arithmetic rounded one operation at a time, then truncated
__device__ int32_t chain_sc(const AncRec &ai, const AncRec &aj, int32_t max_dist_x, int32_t max_dist_y, int32_t bw, float pen_gap)
{
#pragma clang fp contract(off)
	const int64_t dq = (int64_t)ai.qpos - aj.qpos;
	if (dq <= 0 || dq > max_dist_y + ai.len) return INT32_MIN;
	if (ai.sid != aj.sid) return INT32_MIN;
	const int64_t dr = ai.tpos - aj.tpos;
	if (dr <= 0 || dq > max_dist_x + ai.len) return INT32_MIN;
	const int64_t dd = dr > dq ? dr - dq : dq - dr;
	if (dd > bw) return INT32_MIN;
	const int64_t dg = dr < dq ? dr : dq;
	int64_t sc = ai.len < dg ? ai.len : dg;
	if (dd) {
		const float lin = pen_gap * (float)dd;
		const float half = .5f * mb_log2_dev((float)(dd + 1));
		sc -= (int)(lin + half);
	}
	return (int32_t)sc;
}

// mb_chain_bk_end (lchain.c:25-42)
__device__ int64_t chain_bk_end(int32_t max_drop, const U128 *z, const int32_t *f, const int64_t *p, int32_t *t, int64_t k)
{
	int64_t i = (int64_t)z[k].y, end_i = -1, max_i = i;
	int32_t max_s = 0;
	if (i < 0 || t[i] != 0) return i;
	do {
		t[i] = 2;
		end_i = i = p[i];
		const int32_t s = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
		if (s > max_s) max_s = s, max_i = i;
		else if (max_s - s > max_drop) break;
	} while (i >= 0 && t[i] == 0);
	for (i = (int64_t)z[k].y; i >= 0 && i != end_i; i = p[i]) t[i] = 0;
	return max_i;
}

__global__ void __launch_bounds__(64) lchain_kernel(ChainArgs a)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= a.n_reads) return;
	const int64_t o = a.aoff[r], n = a.nanc[r];
	const AncRec *an = a.anc + o;
	int32_t *f = a.f + o, *t = a.t + o, *v = a.v + o;
	int64_t *p = a.p + o;
	U128 *z = a.z + o;
	uint64_t *u = a.ut + o;
	int32_t n_u = 0;
	int64_t n_v = 0;
	ChainOpt co;
	if (a.opts) co = a.opts[a.opt_idx[r]];
	else co.max_dist_x = a.max_dist_x, co.max_dist_y = a.max_dist_y, co.bw = a.bw, co.max_skip = a.max_skip, co.max_iter = a.max_iter,
	     co.min_sc = a.min_sc, co.pen_gap = a.pen_gap;
	const int32_t bw = co.bw, max_drop = co.bw, max_iter = co.max_iter, max_skip = co.max_skip, min_sc = co.min_sc;
	const float pen_gap = co.pen_gap;
	const int32_t mdx = co.max_dist_x < bw ? bw : co.max_dist_x, mdy = co.max_dist_y < bw ? bw : co.max_dist_y;
	a.n_u[r] = 0, a.n_b[r] = 0;
	if (n <= 0) return;
	for (int64_t i = 0; i < n; ++i) t[i] = 0;
	// the DP (lchain.c:170-222)
	int64_t max_ii = -1;
	for (int64_t i = 0; i < n; ++i) {
		const AncRec ai = an[i];
		int64_t max_j = -1, j;
		int32_t max_f = ai.len, n_skip = 0;
		for (j = i - 1; j >= 0 && j >= i - max_iter; --j) {
			const AncRec aj = an[j];
			if (ai.tpos - aj.tpos >= (int64_t)mdx + ai.len) break;
			int32_t sc = chain_sc(ai, aj, mdx, mdy, bw, pen_gap);
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
		const int64_t end_j = j;
		if (max_ii < 0 || ai.tpos - an[max_ii].tpos > (int64_t)mdx + ai.len) {
			int32_t mx = INT32_MIN;
			max_ii = -1;
			for (j = i - 1; j >= end_j && j >= 0; --j)
				if (mx < f[j]) mx = f[j], max_ii = j;
		}
		if (max_ii >= 0 && max_ii < end_j) {
			const int32_t tmp = chain_sc(ai, an[max_ii], mdx, mdy, bw, pen_gap);
			if (tmp != INT32_MIN && max_f < tmp + f[max_ii]) max_f = tmp + f[max_ii], max_j = max_ii;
		}
		f[i] = max_f, p[i] = max_j;
		if (max_ii < 0 || (ai.tpos - an[max_ii].tpos <= (int64_t)mdx + ai.len && f[max_ii] < f[i])) max_ii = i;
	}
	// the backtrack (lchain.c:44-92), in one pass: the reference's counting pass leaves the same state
	int32_t bb[256], be[256];
	int64_t n_z = 0;
	for (int64_t i = 0; i < n; ++i)
		if (f[i] >= min_sc) z[n_z].x = (uint64_t)(int64_t)f[i], z[n_z].y = (uint64_t)i, ++n_z;
	if (n_z == 0) return;
	rs_sort(z, n_z, KeyU128(), bb, be);
	for (int64_t i = 0; i < n; ++i) t[i] = 0;
	for (int64_t k = n_z - 1; k >= 0; --k) {
		if (t[z[k].y] != 0) continue;
		const int64_t n_v0 = n_v, end_i = chain_bk_end(max_drop, z, f, p, t, k);
		int64_t i = (int64_t)z[k].y;
		for (; i != end_i; i = p[i]) v[n_v++] = (int32_t)i, t[i] = 1;
		const int32_t sc = i < 0 ? (int32_t)z[k].x : (int32_t)z[k].x - f[i];
		if (sc >= min_sc && n_v > n_v0) u[n_u++] = (uint64_t)(uint32_t)sc << 32 | (uint64_t)(n_v - n_v0);
		else n_v = n_v0;
	}
	if (n_u == 0) return;
	// compact_a (lchain.c:94-123): each chain first anchor first, chains by their first anchor's
	// concatenated tpos
	AncRec *b0 = a.bb + o;
	U128 *w = z;                                   // z is done with
	for (int64_t i = 0, k = 0; i < n_u; ++i) {
		const int64_t k0 = k, ni = (int32_t)u[i];
		for (int64_t j = 0; j < ni; ++j) b0[k++] = an[v[k0 + (ni - j - 1)]];
	}
	for (int64_t i = 0, k = 0; i < n_u; ++i) {
		const int64_t c = b0[k].sid >> 1;
		w[i].x = (uint64_t)(b0[k].tpos + (int64_t)(a.ctg_off[c] * 2 + a.ctg_len[c] * (uint64_t)(b0[k].sid & 1)));
		w[i].y = (uint64_t)k << 32 | (uint64_t)i;
		k += (int32_t)u[i];
	}
	rs_sort(w, n_u, KeyU128(), bb, be);
	uint64_t *uo = a.u + o;
	AncRec *bo = a.b + o;
	int64_t k = 0;
	for (int64_t i = 0; i < n_u; ++i) {
		const int32_t jj = (int32_t)w[i].y, m = (int32_t)u[jj];
		uo[i] = u[jj];
		const AncRec *src = b0 + (w[i].y >> 32);
		for (int32_t q = 0; q < m; ++q) bo[k + q] = src[q];
		k += m;
	}
	a.n_u[r] = n_u, a.n_b[r] = k;
}

hipError_t launch_lchain(const ChainArgs &a, hipStream_t st)
{
	if (a.n_reads <= 0) return hipSuccess;
	hipLaunchKernelGGL(lchain_kernel, dim3((unsigned)((a.n_reads + 63) / 64)), dim3(64), 0, st, a);
	return hipGetLastError();
}

hipError_t launch_anc_rows(const AncArgs &a, hipStream_t st)
{
	if (a.n_reads <= 0) return hipSuccess;
	hipLaunchKernelGGL(anc_rows_kernel, dim3((unsigned)((a.n_reads + 63) / 64)), dim3(64), 0, st, a);
	return hipGetLastError();
}

hipError_t launch_anc_build(const AncArgs &a, hipStream_t st)
{
	if (a.n_reads <= 0) return hipSuccess;
	hipLaunchKernelGGL(anc_build_kernel, dim3((unsigned)((a.n_reads + 63) / 64)), dim3(64), 0, st, a);
	return hipGetLastError();
}

} // namespace mbg
