// mbg_fm.hip -- FM-index kernels for gfx950: k-mer interval cache, rank/extend probes, two-pass
// SMEM seeding (a persistent lane-per-task state machine), SA locate, target extraction.
//
// Seeding restates mb_seed_intv_batch (seed.c:56-103) over mb_bwt_smem_batch's six stages
// (bwt.c:356-463).  The CPU interleaves 50 reads through a queue so their DRAM misses overlap; here
// every lane owns one task (a read in pass 1, one eligible SMEM in pass 2), the whole batch is in
// flight at once, and a lane that finishes pulls the next task from a global counter, so waves stay
// full while read lengths and stage counts vary.  Each loop trip does at most one dependent memory
// step (one occ-block rank or one 32-byte cache lookup), the unit the reference also interleaves.
#include "mbg_internal.h"

namespace mbg {

// ---------------------------------------------------------------------------------------------
// k-mer interval cache (mb_bwt_count_kmer / mb_bwt_cache, bwt.c:552-591)
// key: the pattern's last base in the top two bits, its first base in the lowest two bits
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kmer_cache_kernel(Idx ix, Sai *pre, int len, uint64_t n)
{
	const uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (z >= n) return;
	Sai p = set_intv(ix, (int)(z >> (2 * (len - 1)) & 3));
	for (int d = len - 2; d >= 0; --d)
		p = extend1(ix, p, 1, (int)(z >> (2 * d) & 3));
	p.info = 0;
	pre[z] = p;
}

__global__ void __launch_bounds__(256) rank2a_kernel(Idx ix, int64_t n, const uint64_t *k, const uint64_t *l, uint64_t *ck, uint64_t *cl)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	uint64_t a[4], b[4];
	rank2a(ix, k[i], l[i], a, b);
	for (int c = 0; c < 4; ++c) ck[i * 4 + c] = a[c], cl[i * 4 + c] = b[c];
}

__global__ void __launch_bounds__(256) extend_kernel(Idx ix, int64_t n, const Sai *ik, const uint8_t *is_back, Sai *ok)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	Sai o[4];
	extend4(ix, ik[i], is_back[i] ? 1 : 0, o);
	for (int c = 0; c < 4; ++c) ok[i * 4 + c] = o[c];
}

// ---------------------------------------------------------------------------------------------
// SMEM state machine
// ---------------------------------------------------------------------------------------------
enum : int { ACT_DONE = 0, ACT_LOOKUP = 1, ACT_BACK = 2, ACT_FWD = 3 };

struct Lane {
	const uint8_t *q;
	int32_t en, min_len, x, i, stage;
	uint32_t kmer, rid;
	uint64_t min_occ;
	Sai p;
};

__device__ __forceinline__ void emit(const SeedArgs &a, const Lane &s, uint32_t pass)
{
	const unsigned long long slot = atomicAdd((unsigned long long*)a.n_out, 1ULL);
	if ((int64_t)slot < a.cap) {
		SeedRec r;
		r.x0 = s.p.x0; r.x1 = s.p.x1; r.size = s.p.size;
		r.info = (uint64_t)(uint32_t)s.x << 32 | (uint64_t)(uint32_t)s.i;
		r.rid = s.rid; r.pass = pass;
		a.out[slot] = r;
	}
}

// Cheap transitions until the lane needs memory (bwt.c:381-459 minus the rank/lookup steps).
__device__ __forceinline__ int control(const Idx &ix, const SeedArgs &a, Lane &s, uint32_t pass)
{
	for (;;) {
		if (s.stage == 1) {
			if (s.en - s.x < s.min_len) return ACT_DONE;
			int xn = -1;
			for (int i = s.x; i < s.x + s.min_len; ++i)
				if (s.q[i] > 3) xn = i;
			if (xn >= 0) { s.x = xn + 1; continue; }
			s.i = s.x + s.min_len - 1;
			if (ix.pre_len > 0 && s.min_len >= ix.pre_len) {
				uint32_t z = 0;
				for (int l = 0; l < ix.pre_len; ++l, --s.i) z = z << 2 | s.q[s.i];
				s.kmer = z;
				s.stage = 2;
				return ACT_LOOKUP;
			}
			s.p = set_intv(ix, s.q[s.i]);
			--s.i;
			s.stage = 3;
			continue;
		} else if (s.stage == 3) {
			if (s.i < s.x) { s.i = s.x + s.min_len; s.stage = 4; continue; }
			return ACT_BACK;
		} else if (s.stage == 4) {
			if (s.i == s.en) { emit(a, s, pass); return ACT_DONE; }
			if (s.q[s.i] < 4) return ACT_FWD;
			emit(a, s, pass);                           // N: save, skip it, back to stage 1
			s.x = s.i + 1;
			s.stage = 1;
			continue;
		} else if (s.stage == 6) {
			if (s.i < s.x + 1) { s.x = s.i + 1; s.stage = 1; continue; }
			return ACT_BACK;
		} else { // stages 2 and 5 are entered only through a lookup
			return ACT_LOOKUP;
		}
	}
}

template <int PASS>
__global__ void __launch_bounds__(256) smem_kernel(Idx ix, SeedArgs a)
{
	const uint32_t pass = PASS;
	Lane s;
	int act = ACT_DONE;
	int nblk = 0, nlook = 0, next_ = 0;
	uint32_t n_tasks = a.n_tasks;
	if (a.n_tasks_dev) {
		const int64_t nt = (int64_t)*a.n_tasks_dev;
		n_tasks = (uint32_t)(nt < a.task_cap ? nt : a.task_cap);
	}
	for (;;) {
		const uint32_t t = wave_fetch(a.next, act == ACT_DONE);
		if (act == ACT_DONE) {
			if (t >= n_tasks) {
				if (a.stats) {
					atomicAdd(&a.stats[0], (unsigned long long)nblk);
					atomicAdd(&a.stats[1], (unsigned long long)nlook);
					atomicAdd(&a.stats[2], (unsigned long long)next_);
				}
				break;
			}
			if (pass == 1) {
				s.rid = t;
				s.q = a.seq + a.off[t];
				s.x = 0; s.en = a.len[t];
				s.min_len = a.min_len; s.min_occ = 1;
			} else {
				const SeedTask tk = a.tasks[t];
				s.rid = tk.rid;
				s.q = a.seq + a.off[tk.rid];
				s.x = tk.st; s.en = tk.en;
				s.min_len = tk.min_len; s.min_occ = (uint64_t)(int64_t)tk.min_occ;
			}
			s.stage = 1;
		}
		act = control(ix, a, s, pass);
		if (act == ACT_DONE) continue;
		if (act == ACT_LOOKUP) {                        // stages 2/5 (bwt.c:401-413)
			s.p = ix.pre[s.kmer];
			++nlook;
			if (s.p.size < s.min_occ) {
				s.i += ix.pre_len;
				s.p = set_intv(ix, s.q[s.i]);
				--s.i;
			}
			++s.stage;
		} else if (act == ACT_BACK) {                   // stages 3/6: se_one_step_back (bwt.c:339-354)
			const Sai o = extend1(ix, s.p, 1, s.q[s.i], &nblk);
			++next_;
			if (o.size < s.min_occ) { s.x = s.i + 1; s.stage = 1; }
			else { s.p = o; --s.i; }
		} else {                                        // stage 4 forward step (bwt.c:421-453)
			const Sai o = extend1(ix, s.p, 0, 3 - (int)s.q[s.i], &nblk);
			++next_;
			if (o.size >= s.min_occ) { s.p = o; ++s.i; }
			else {
				emit(a, s, pass);
				if (ix.pre_len > 0 && s.i - s.x - 1 >= ix.pre_len) {
					uint32_t z = 0;
					for (int l = 0; l < ix.pre_len; ++l, --s.i) z = z << 2 | s.q[s.i];
					s.kmer = z;
					s.stage = 5;
				} else {
					s.p = set_intv(ix, s.q[s.i]);
					--s.i;
					s.stage = 6;
				}
			}
		}
	}
}

// pass-2 tasks (seed.c:80-98): every pass-1 SMEM with length >= 2*min_len and size <= max_sub_occ
__global__ void __launch_bounds__(256) subtask_kernel(const SeedRec *out, const int64_t *n_out, int64_t cap, int32_t min_len,
                                                     int32_t max_sub_occ, SeedTask *tasks, uint32_t *n_tasks, int64_t task_cap)
{
	const int64_t n = *n_out < cap ? *n_out : cap;
	for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
		const SeedRec r = out[j];
		const uint32_t st = (uint32_t)(r.info >> 32), en = (uint32_t)r.info;
		if (en - st < (uint32_t)(min_len * 2) || r.size > (uint64_t)(int64_t)max_sub_occ) continue;
		const uint32_t k = atomicAdd(n_tasks, 1u);
		if ((int64_t)k >= task_cap) continue;
		SeedTask t;
		t.rid = r.rid;
		t.st = (int32_t)st; t.en = (int32_t)en;
		t.min_len = (int32_t)(en - st) / 2 > min_len ? (int32_t)(en - st) / 2 : min_len;
		t.min_occ = (int32_t)r.size + 1;
		tasks[k] = t;
	}
}

// ---------------------------------------------------------------------------------------------
// SA locate (mb_bwt_sa / mb_bwt_sa_batch, bwt.c:499-541): LF-walk each row to the nearest sampled
// row; lanes refill from a counter because walk lengths are geometric (mean 2^sa_bit).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) sa_kernel(Idx ix, int64_t n, const int64_t *n_dev, uint64_t *x, uint32_t *next,
                                                unsigned long long *stats)
{
	const uint64_t mask = (1ULL << ix.sa_bit) - 1;
	if (n_dev) { const int64_t nd = *n_dev; n = nd < n ? nd : n; }
	bool have = false;
	uint32_t idx = 0;
	uint64_t k = 0, steps = 0, tot = 0;
	for (;;) {
		const uint32_t t = wave_fetch(next, !have);
		if (!have) {
			if ((int64_t)t >= n) break;
			idx = t; k = x[t]; steps = 0; have = true;
		}
		if ((k & mask) == 0) {
			x[idx] = steps + ix.sa[k >> ix.sa_bit];
			tot += steps;
			have = false;
			continue;
		}
		k = lf(ix, k);
		++steps;
	}
	if (stats) atomicAdd(stats, (unsigned long long)tot);
}

// ---------------------------------------------------------------------------------------------
// Target extraction (l2b_getseq + l2b_getambi, l2bit.c:52-121): one wave per window
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(64) getseq_kernel(L2b lb, int64_t n, const int64_t *tid, const int64_t *st_, const int64_t *en_,
                                                   const int64_t *out_off, uint8_t *out, int64_t *len)
{
	const int64_t w = blockIdx.x;
	if (w >= n) return;
	const int lane = threadIdx.x;
	const int64_t t = tid[w];
	if (t < 0 || (uint64_t)t >= lb.n_ctg) { if (lane == 0) len[w] = -1; return; }
	const uint64_t coff = lb.ctg_off[t], clen = lb.ctg_len[t];
	int64_t st = st_[w], en = en_[w];
	if (st < 0) st = 0;
	if (en > (int64_t)clen) en = (int64_t)clen;
	uint8_t *o = out + out_off[w];
	const int64_t gst = (int64_t)coff + st, gen = (int64_t)coff + en;
	for (int64_t i = gst + lane; i < gen; i += 64)
		o[i - gst] = (uint8_t)(lb.pac[i >> 5] >> ((i & 31) << 1) & 3);
	if (lane == 0) len[w] = en - st;
	if (st >= en || lb.n_ambi == 0) return;
	// l2b_getambi: first run with en > gst, first run with st >= gen
	int64_t lo = 0, hi = (int64_t)lb.n_ambi;
	while (lo < hi) { int64_t mid = (lo + hi) / 2; if ((int64_t)lb.ambi[mid * 2 + 1] > gst) hi = mid; else lo = mid + 1; }
	const int64_t i_st = lo;
	lo = i_st, hi = (int64_t)lb.n_ambi;
	while (lo < hi) { int64_t mid = (lo + hi) / 2; if ((int64_t)lb.ambi[mid * 2] >= gen) hi = mid; else lo = mid + 1; }
	const int64_t i_en = lo;
	__syncthreads();
	for (int64_t a = i_st; a < i_en; ++a) {
		int64_t s = (int64_t)lb.ambi[a * 2], e = (int64_t)lb.ambi[a * 2 + 1];
		if (s < gst) s = gst;
		if (e > gen) e = gen;
		for (int64_t i = s + lane; i < e; i += 64) o[i - gst] = 4;
	}
}

// plain streaming copy (HBM byte-counter calibration)
__global__ void __launch_bounds__(256) stream_copy_kernel(uint4 *dst, const uint4 *src, int64_t n16)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
		dst[i] = src[i];
}

hipError_t launch_stream_copy(void *dst, const void *src, int64_t n16, unsigned grid, hipStream_t st)
{
	hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, st, (uint4*)dst, (const uint4*)src, n16);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_kmer_cache(const Idx &ix, Sai *pre, int len, hipStream_t st)
{
	const uint64_t n = 1ULL << (2 * len);
	hipLaunchKernelGGL(kmer_cache_kernel, dim3(blocks_for((int64_t)n, 256)), dim3(256), 0, st, ix, pre, len, n);
	return hipGetLastError();
}

hipError_t launch_rank2a(const Idx &ix, int64_t n, const uint64_t *k, const uint64_t *l, uint64_t *ck, uint64_t *cl, hipStream_t st)
{
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(rank2a_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, ix, n, k, l, ck, cl);
	return hipGetLastError();
}

hipError_t launch_extend(const Idx &ix, int64_t n, const Sai *ik, const uint8_t *is_back, Sai *ok, hipStream_t st)
{
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(extend_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, ix, n, ik, is_back, ok);
	return hipGetLastError();
}

hipError_t launch_smem(const Idx &ix, const SeedArgs &a, uint32_t pass, unsigned grid, hipStream_t st)
{
	if (pass == 1) hipLaunchKernelGGL(smem_kernel<1>, dim3(grid), dim3(256), 0, st, ix, a);
	else hipLaunchKernelGGL(smem_kernel<2>, dim3(grid), dim3(256), 0, st, ix, a);
	return hipGetLastError();
}

hipError_t launch_subtask(const SeedRec *out, const int64_t *n_out, int64_t cap, int32_t min_len, int32_t max_sub_occ,
                          SeedTask *tasks, uint32_t *n_tasks, int64_t task_cap, unsigned grid, hipStream_t st)
{
	hipLaunchKernelGGL(subtask_kernel, dim3(grid), dim3(256), 0, st, out, n_out, cap, min_len, max_sub_occ, tasks, n_tasks, task_cap);
	return hipGetLastError();
}

hipError_t launch_sa(const Idx &ix, int64_t n, const int64_t *n_dev, uint64_t *x, uint32_t *next, unsigned long long *stats,
                     unsigned grid, hipStream_t st)
{
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(sa_kernel, dim3(grid), dim3(256), 0, st, ix, n, n_dev, x, next, stats);
	return hipGetLastError();
}

hipError_t launch_getseq(const L2b &lb, int64_t n, const int64_t *tid, const int64_t *st_, const int64_t *en_, const int64_t *out_off,
                         uint8_t *out, int64_t *len, hipStream_t st)
{
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(getseq_kernel, dim3((unsigned)n), dim3(64), 0, st, lb, n, tid, st_, en_, out_off, out, len);
	return hipGetLastError();
}

} // namespace mbg
