// mbg_anchor.hip -- mb_anchor on gfx950 (seed.c:109-339, no methylation conversion): a read's SMEM
// intervals become contig-local anchors, in the reference's order.
//
//   anc_rows_kernel      one thread per read: mb_seed_sort_dedup, then the rows to locate for each
//                        group of intervals sharing (x[0], size): all of a group of size <= max_occ,
//                        else max_occ evenly stepped ones (seed.c:304-315)
//   sa_kernel            (mbg_fm.hip) the rows' SA positions, in place
//   anc_build_kernel     one thread per read: each position of each interval of its group mapped to a
//                        contig and strand (seed.c:176-210), the anchors sorted by concatenated tpos,
//                        made contig-local, deduplicated (seed.c:140-166) and compacted
//
// Both sorts are klib's radix sort (ksort.h:108-161): insertion sort up to 64 elements, else an
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
		if (m + ns > cap) { a.nu[r] = -1; return; }
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
