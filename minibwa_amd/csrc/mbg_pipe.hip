// mbg_pipe.hip -- device-resident seed -> locate -> extend pass over a batch of reads.
//
// After two-pass seeding the step samples anchor rows, locates them, and builds one extension job
// per read end from the read's longest SMEM, with the window and fast-path rules the reference's
// short-read aligner applies to a chain:
//   * SA rows per interval (seed.c:304-315): all rows when size <= max_occ, else max_occ rows with
//     step (size - j) / (max_occ - n);
//   * position -> contig/strand (l2b_intv2cid, l2bit.c:9-34);
//   * left/right windows (mb_align1 short-read mode, align.c:588-596, 640-653, 732-742): the query
//     end beyond the anchor plus `l` target bases, l += (l*a + end_bonus - q) / e when positive;
//   * mb_align_pair (align.c:329-378): ungapped extension first (accepted with at most 3
//     mismatches when it reaches the end within end_bonus of its maximum), otherwise a ksw job
//     with the band capped by max_bw_from_mm when the shorter side is < 100.
// Only jobs that fail the fast path reach the DP kernel, as in the reference.
#include "mbg_internal.h"

namespace mbg {

// number of SA rows mb_anchor locates for an interval of `size` rows (seed.c:304-315)
__device__ __forceinline__ int64_t n_anchor_rows(uint64_t size, int32_t max_occ)
{
	if ((int64_t)size <= max_occ) return (int64_t)size;
	int64_t cnt = 0;
	for (int64_t j = 0, k = 0; j < (int64_t)size && k < max_occ; ++k) {
		int64_t step = ((int64_t)size - j) / (max_occ - k);
		if (step < 1) step = 1;
		j += step, ++cnt;
	}
	return cnt;
}

// rows per record (zero past the record count), the input of the exclusive scan
__global__ void __launch_bounds__(256) row_counts_kernel(const SeedRec *recs, const int64_t *n_rec_dev, int64_t rec_cap, int32_t max_occ, uint32_t *cnt)
{
	int64_t n = *n_rec_dev;
	if (n > rec_cap) n = rec_cap;
	for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rec_cap; r += (int64_t)gridDim.x * blockDim.x)
		cnt[r] = r < n ? (uint32_t)n_anchor_rows(recs[r].size, max_occ) : 0;
}

// SA rows of each interval at its scanned offset; first_slot[rec] = slot of its first row
__global__ void __launch_bounds__(256) anchor_rows_kernel(const SeedRec *recs, const int64_t *n_rec_dev, int64_t rec_cap, int32_t max_occ,
                                                         const uint32_t *cnt, const uint32_t *off, uint64_t *rows, uint32_t *row_rec,
                                                         uint32_t *first_slot, unsigned long long *n_rows, int64_t row_cap)
{
	int64_t n = *n_rec_dev;
	if (n > rec_cap) n = rec_cap;
	if (blockIdx.x == 0 && threadIdx.x == 0) *n_rows = rec_cap > 0 ? (unsigned long long)off[rec_cap - 1] + cnt[rec_cap - 1] : 0;
	for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
		const SeedRec s = recs[r];
		const int64_t c = cnt[r], base = off[r];
		first_slot[r] = c > 0 && base < row_cap ? (uint32_t)base : 0xffffffffu;
		if (base + c > row_cap) continue;
		if ((int64_t)s.size <= max_occ) {
			for (int64_t j = 0; j < c; ++j) rows[base + j] = s.x0 + j, row_rec[base + j] = (uint32_t)r;
		} else {
			int64_t o = 0;
			for (int64_t j = 0, k = 0; j < (int64_t)s.size && k < max_occ; ++k) {
				int64_t step = ((int64_t)s.size - j) / (max_occ - k);
				if (step < 1) step = 1;
				rows[base + o] = s.x0 + j, row_rec[base + o] = (uint32_t)r, ++o;
				j += step;
			}
		}
	}
}

// longest pass-1 SMEM of each read (ties: smallest query start, then first record)
__global__ void __launch_bounds__(256) best_seed_kernel(const SeedRec *recs, const int64_t *n_rec_dev, int64_t rec_cap, unsigned long long *best)
{
	int64_t n = *n_rec_dev;
	if (n > rec_cap) n = rec_cap;
	for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
		const SeedRec s = recs[r];
		if (s.pass != 1 || s.size == 0) continue;
		const uint32_t qs = (uint32_t)(s.info >> 32), qe = (uint32_t)s.info;
		const uint64_t len = qe - qs;
		const uint64_t key = (len < 0xffff ? len : 0xffff) << 48 | (uint64_t)(0xffffu - (qs < 0xffff ? qs : 0xffff)) << 32 | (uint64_t)(0xffffffffu - (uint32_t)r);
		atomicMax(&best[s.rid], (unsigned long long)key);
	}
}

struct JobSink {
	DpJob *jobs;
	JobMeta *meta;
	uint8_t *qbuf, *tbuf;
	int64_t job_cap, qcap, tcap;
	unsigned long long *ctr;   // [0] jobs [1] qbytes [2] tbytes [3] fast-path accepts [4] anchored reads [5] cigar slots
};

// target bases of concatenated coordinates [gst, gen) read forward or reversed; N runs read as 4
// (l2b_getseq + l2b_getambi, l2bit.c:52-121)
struct Window {
	const uint64_t *pac = nullptr, *ambi = nullptr;
	int64_t gst = 0, gen = 0;
	bool reverse = false;
	int64_t a0 = 0, a1 = 0;    // N runs overlapping the window
	__device__ void init(const L2b &l, int64_t s, int64_t e, bool rev)
	{
		pac = l.pac, ambi = l.ambi, gst = s, gen = e, reverse = rev, a0 = a1 = 0;
		if (l.n_ambi == 0 || e <= s) return;
		int64_t lo = 0, hi = (int64_t)l.n_ambi;
		while (lo < hi) { int64_t mid = (lo + hi) / 2; if ((int64_t)l.ambi[mid * 2 + 1] > s) hi = mid; else lo = mid + 1; }
		a0 = lo, a1 = lo;
		while (a1 < (int64_t)l.n_ambi && (int64_t)l.ambi[a1 * 2] < e) ++a1;
	}
	__device__ uint8_t operator()(int64_t i) const
	{
		const int64_t p = reverse ? gen - 1 - i : gst + i;
		for (int64_t a = a0; a < a1; ++a)
			if ((int64_t)ambi[a * 2] <= p && p < (int64_t)ambi[a * 2 + 1]) return 4;
		return (uint8_t)(pac[p >> 5] >> ((p & 31) << 1) & 3);
	}
	__device__ bool plain() const { return a0 == a1; }   // no N run overlaps the window
};

// the bases of a Window without N runs in order, one 64-bit pac load per 32 bases
struct WinCursor {
	const uint64_t *pac;
	int64_t p, wi = -1;
	int step;
	uint64_t w = 0;
	__device__ explicit WinCursor(const Window &t) : pac(t.pac), p(t.reverse ? t.gen - 1 : t.gst), step(t.reverse ? -1 : 1) {}
	__device__ uint8_t next()
	{
		const int64_t k = p >> 5;
		if (k != wi) wi = k, w = pac[k];
		const uint8_t c = (uint8_t)(w >> ((p & 31) << 1) & 3);
		p += step;
		return c;
	}
};

// the read as aligned: forward, or reverse-complemented with N kept (align.c:902-903), then a
// [start, start+n) slice, read forward or reversed
struct QView {
	const uint8_t *r;
	int32_t qlen, start;
	bool rc, reverse;
	int32_t n;
	__device__ uint8_t operator()(int32_t i) const
	{
		const int32_t j = start + (reverse ? n - 1 - i : i);      // index on the oriented read
		if (!rc) return r[j];
		const uint8_t c = r[qlen - 1 - j];
		return c < 4 ? (uint8_t)(3 - c) : (uint8_t)4;
	}
};

__device__ __forceinline__ int max_bw_from_mm(const MapOpt &o, int mm) // align.c:321-327
{
	const int x = mm * (o.a + o.b);
	int max2 = 0, max1 = 0;
	if (x >= o.q + o.e) max1 = (x - o.q + o.e - 1) / o.e;
	if (x >= o.q2 + o.e2) max2 = (x - o.q2 + o.e2 - 1) / o.e2;
	return max1 > max2 ? max1 : max2;
}

// one extension (mb_align_pair with KSW_EZ_EXTZ_ONLY, align.c:329-378): the ungapped fast path,
// else the band (capped by max_bw_from_mm when the shorter side is < 100) of a DP job
struct ExtPlan {
	bool job = false, fast = false;
	int qlen = 0, tlen = 0, w = 0, flag = 0, side = 0;
	int64_t tpos = 0;
	QView q{};
	Window t{};
};

// ungapped score of the extension: n_mm (starting at -1, as in align.c:337), final score and max
template <class TB>
__device__ __forceinline__ void ungapped(const int8_t *mat, const QView &q, int qlen, TB &&tb, int &n_mm, int32_t &score, int32_t &mx)
{
	for (int j = 0; j < qlen; ++j) {
		const uint8_t tj = tb(j), qj = q(j);
		const int32_t sc = mat[tj * 5 + qj];
		score += sc;
		n_mm += (tj > 3 || qj > 3 || sc < 0);
		if (mx < score) mx = score;
	}
}

// mat: the 5x5 scoring matrix staged in LDS
__device__ __forceinline__ void plan_ext(const MapOpt &o, const int8_t *mat, ExtPlan &p)
{
	const int qlen = p.q.n, tlen = p.tlen;
	int w = o.ksw_w, n_mm = -1;
	p.qlen = qlen;
	if (tlen >= qlen) {
		int32_t score = 0, mx = 0;
		if (p.t.plain()) {
			WinCursor c(p.t);
			ungapped(mat, p.q, qlen, [&](int) { return c.next(); }, n_mm, score, mx);
		} else {
			ungapped(mat, p.q, qlen, [&](int j) { return p.t(j); }, n_mm, score, mx);
		}
		if (n_mm <= 2 && score + o.end_bonus >= mx) { p.fast = true; return; }
	}
	if (n_mm >= 0 && (qlen < tlen ? qlen : tlen) < 100) {
		const int mbw = max_bw_from_mm(o, n_mm);
		if (w > mbw + 4) w = mbw + 4;
	}
	p.w = w;
	p.job = true;
}

__device__ __forceinline__ void write_job(const MapOpt &o, const JobSink &js, const ExtPlan &p, uint64_t j, uint64_t qo, uint64_t to,
                                          uint64_t co, uint32_t rid, int rev, int64_t tid)
{
	if ((int64_t)j >= js.job_cap || (int64_t)(qo + p.qlen) > js.qcap || (int64_t)(to + p.tlen) > js.tcap) {
		// a read longer than the step's max_len: the job, already counted, becomes an empty one (the DP
		// kernels skip qlen 0) and the overflow is counted, so that the step reports it
		if ((int64_t)j < js.job_cap) {
			DpJob d{};
			d.cigar_off = (int64_t)co;
			js.jobs[j] = d;
			JobMeta m{};
			m.rid = rid; m.side = (int8_t)p.side; m.rev = (int8_t)rev; m.tid = (int32_t)tid; m.tpos = p.tpos;
			js.meta[j] = m;
		}
		atomicAdd(&js.ctr[15], 1ull);
		return;
	}
	for (int i = 0; i < p.qlen; ++i) js.qbuf[qo + i] = p.q(i);
	if (p.t.plain()) {
		WinCursor c(p.t);
		for (int i = 0; i < p.tlen; ++i) js.tbuf[to + i] = c.next();
	} else {
		for (int i = 0; i < p.tlen; ++i) js.tbuf[to + i] = p.t(i);
	}
	DpJob d;
	d.qoff = (int64_t)qo; d.toff = (int64_t)to; d.qlen = p.qlen; d.tlen = p.tlen;
	d.q = (int8_t)o.q; d.e = (int8_t)o.e; d.q2 = (int8_t)o.q2; d.e2 = (int8_t)o.e2;
	d.w = p.w; d.zdrop = o.zdrop; d.end_bonus = o.end_bonus; d.flag = p.flag; d.cigar_off = (int64_t)co;
	js.jobs[j] = d;
	JobMeta m;
	m.rid = rid; m.side = (int8_t)p.side; m.rev = (int8_t)rev; m.pad = 0; m.tid = (int32_t)tid; m.tpos = p.tpos;
	js.meta[j] = m;
}

__device__ __forceinline__ uint64_t wave_bcast_u64(uint64_t v)
{
	const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, 0), hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), 0);
	return (uint64_t)hi << 32 | lo;
}

// per read: anchor = longest SMEM at its first located row; left and right extension jobs.  Output
// positions come from wave prefix sums and one atomic per counter per wave.
__global__ void __launch_bounds__(128) ext_jobs_kernel(L2b lb, MapOpt o, int32_t n_reads, const int64_t *off, const int32_t *len,
                                                      const uint8_t *seq, const SeedRec *recs, const uint32_t *first_slot,
                                                      const uint64_t *pos, const unsigned long long *best, JobSink js)
{
	__shared__ int8_t mat[28];
	if (threadIdx.x < 28) mat[threadIdx.x] = o.mat[threadIdx.x];
	__syncthreads();
	const int32_t rid = blockIdx.x * blockDim.x + threadIdx.x;
	ExtPlan L, R;
	bool anchored = false;
	int rev = 0;
	int64_t tid = -1;
	if (rid < n_reads) {
		const uint64_t key = best[rid];
		const uint32_t rec = 0xffffffffu - (uint32_t)key;
		const uint32_t fs = key ? first_slot[rec] : 0xffffffffu;
		if (fs != 0xffffffffu) {
			const SeedRec s = recs[rec];
			const int32_t qlen = len[rid];
			const int32_t qs0 = (int32_t)(s.info >> 32), qe0 = (int32_t)s.info, alen = qe0 - qs0;
			// l2b_intv2cid (l2bit.c:9-34)
			const uint64_t st = pos[fs], en = st + (uint64_t)alen, tot = lb.tot_len;
			if (!(en > tot * 2 || (st < tot && tot < en))) {
				rev = st >= tot;
				const int64_t sp = (int64_t)(st < tot ? st : tot * 2 - en);
				int64_t lo = 0, hi = (int64_t)lb.n_ctg;
				while (lo < hi) {
					const int64_t mid = (lo + hi) / 2;
					const int64_t co = (int64_t)lb.ctg_off[mid], cl = (int64_t)lb.ctg_len[mid];
					if (co <= sp && sp < co + cl) { tid = sp + alen <= co + cl ? mid : -1; break; }
					else if (sp < co) hi = mid; else lo = mid + 1;
				}
				if (tid >= 0) {
					anchored = true;
					const int64_t coff = (int64_t)lb.ctg_off[tid], clen = (int64_t)lb.ctg_len[tid];
					const int64_t ts = sp - coff, te = ts + alen;
					const int32_t qs = rev ? qlen - qe0 : qs0, qe = rev ? qlen - qs0 : qe0;   // on the oriented read
					const uint8_t *r = seq + off[rid];
					if (qs > 0 && ts > 0) { // left: reversed query prefix against the reversed target before ts
						int64_t l = qs;
						l += l * o.a + o.end_bonus > o.q ? (l * o.a + o.end_bonus - o.q) / o.e : 0;
						const int64_t ts0 = ts - l > 0 ? ts - l : 0;
						L.q = QView{r, qlen, 0, rev != 0, true, qs};
						L.t.init(lb, coff + ts0, coff + ts, true);
						L.tlen = (int)(ts - ts0), L.flag = 0x40 | 0x02 | 0x80, L.side = 0, L.tpos = ts;
						plan_ext(o, mat, L);
					}
					if (qe < qlen && te < clen) { // right extension
						int64_t l = qlen - qe;
						l += l * o.a + o.end_bonus > o.q ? (l * o.a + o.end_bonus - o.q) / o.e : 0;
						const int64_t te0 = te + l < clen ? te + l : clen;
						R.q = QView{r, qlen, qe, rev != 0, false, qlen - qe};
						R.t.init(lb, coff + te, coff + te0, false);
						R.tlen = (int)(te0 - te), R.flag = 0x40, R.side = 1, R.tpos = te;
						plan_ext(o, mat, R);
					}
				}
			}
		}
	}
	// wave-aggregated output positions
	const uint32_t nj = (uint32_t)L.job + (uint32_t)R.job;
	const uint32_t nq = (L.job ? L.qlen : 0) + (R.job ? R.qlen : 0), nt = (L.job ? L.tlen : 0) + (R.job ? R.tlen : 0);
	const uint32_t nstat = ((uint32_t)L.fast + (uint32_t)R.fast) | (uint32_t)anchored << 16;
	uint32_t tj, tq, tt, ts;
	const uint32_t ej = wave_excl_scan(nj, &tj), eq = wave_excl_scan(nq, &tq), et = wave_excl_scan(nt, &tt);
	wave_excl_scan(nstat, &ts);
	uint64_t bj = 0, bq = 0, bt = 0, bc = 0;
	if (__lane_id() == 0) {
		if (tj) {
			bj = atomicAdd(&js.ctr[0], (unsigned long long)tj);
			bq = atomicAdd(&js.ctr[1], (unsigned long long)tq);
			bt = atomicAdd(&js.ctr[2], (unsigned long long)tt);
			bc = atomicAdd(&js.ctr[5], (unsigned long long)tq + tt);
		}
		if (ts & 0xffffu) atomicAdd(&js.ctr[3], (unsigned long long)(ts & 0xffffu));
		if (ts >> 16) atomicAdd(&js.ctr[4], (unsigned long long)(ts >> 16));
	}
	if (tj == 0) return;
	bj = wave_bcast_u64(bj), bq = wave_bcast_u64(bq), bt = wave_bcast_u64(bt), bc = wave_bcast_u64(bc);
	uint64_t j = bj + ej, qo = bq + eq, to = bt + et, co = bc + eq + et;
	if (L.job) {
		write_job(o, js, L, j, qo, to, co, (uint32_t)rid, rev, tid);
		++j, qo += L.qlen, to += L.tlen, co += L.qlen + L.tlen;
	}
	if (R.job) write_job(o, js, R, j, qo, to, co, (uint32_t)rid, rev, tid);
}

hipError_t launch_row_counts(const SeedRec *recs, const int64_t *n_rec, int64_t rec_cap, int32_t max_occ, uint32_t *cnt, hipStream_t st)
{
	hipLaunchKernelGGL(row_counts_kernel, dim3(2048), dim3(256), 0, st, recs, n_rec, rec_cap, max_occ, cnt);
	return hipGetLastError();
}

hipError_t launch_anchor_rows(const SeedRec *recs, const int64_t *n_rec, int64_t rec_cap, int32_t max_occ, const uint32_t *cnt,
                              const uint32_t *off, uint64_t *rows, uint32_t *row_rec, uint32_t *first_slot,
                              unsigned long long *n_rows, int64_t row_cap, hipStream_t st)
{
	hipLaunchKernelGGL(anchor_rows_kernel, dim3(2048), dim3(256), 0, st, recs, n_rec, rec_cap, max_occ, cnt, off, rows, row_rec,
	                   first_slot, n_rows, row_cap);
	return hipGetLastError();
}

hipError_t launch_best_seed(const SeedRec *recs, const int64_t *n_rec, int64_t rec_cap, unsigned long long *best, unsigned grid, hipStream_t st)
{
	hipLaunchKernelGGL(best_seed_kernel, dim3(grid), dim3(256), 0, st, recs, n_rec, rec_cap, best);
	return hipGetLastError();
}

hipError_t launch_ext_jobs(const L2b &lb, const MapOpt &o, int32_t n_reads, const int64_t *off, const int32_t *len, const uint8_t *seq,
                           const SeedRec *recs, const uint32_t *first_slot, const uint64_t *pos, const unsigned long long *best,
                           DpJob *jobs, JobMeta *meta, uint8_t *qbuf, uint8_t *tbuf, int64_t job_cap, int64_t qcap, int64_t tcap,
                           unsigned long long *ctr, hipStream_t st)
{
	if (n_reads <= 0) return hipSuccess;
	JobSink js{jobs, meta, qbuf, tbuf, job_cap, qcap, tcap, ctr};
	hipLaunchKernelGGL(ext_jobs_kernel, dim3((unsigned)((n_reads + 127) / 128)), dim3(128), 0, st, lb, o, n_reads, off, len, seq, recs,
	                   first_slot, pos, best, js);
	return hipGetLastError();
}

} // namespace mbg
