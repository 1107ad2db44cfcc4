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
	for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < n; z += (uint64_t)gridDim.x * blockDim.x) {
		Sai p = set_intv(ix, (int)(z >> (2 * (len - 1)) & 3));
		for (int d = len - 2; d >= 0; --d)
			p = extend1(ix, p, 1, (int)(z >> (2 * d) & 3));
		p.info = 0;
		pre[z] = p;
	}
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
//
// Each lane copies its task's bases into LDS as 2-bit codes when the window is N-free and at most
// kQWords*16 long, so the per-step base lookups stay off the memory path; other windows read their
// bases from global memory.  Backward and forward steps share one extension call site, so a wave
// whose lanes sit in different stages still issues one rank per loop trip.
// ---------------------------------------------------------------------------------------------
enum : int { ACT_DONE = 0, ACT_LOOKUP = 1, ACT_BACK = 2, ACT_FWD = 3 };
constexpr int kSmemBlock = 256;
constexpr int kQWords = kPackWords;    // up to 384 bases per lane in LDS

struct Lane {
	const uint8_t *q;                  // task bases in global memory (absolute positions)
	uint32_t task, nemit;              // output slot owner and records emitted so far
	int32_t qbase;                     // absolute position of LDS base 0
	int32_t en, min_len, x, i, stage;
	uint32_t kmer, rid;
	uint64_t min_occ;
	uint64_t px0, px1, psize;          // current interval
	bool gq;                           // bases read from global memory
	int32_t stop;                      // segmented pass 1: the first chain value this task does not start from
	int32_t k, K;                      // fix-up: the segment it is at, and the read's segment count
	uint32_t nch;                      // segment task: chain values recorded
};

// stage of a task that gave up (its output slots are full): the read falls back to a serial task
constexpr int kStageAbandon = 7;
constexpr uint32_t kNoConv = 0xffffffffu;

__device__ __forceinline__ int qget(const Lane &s, const uint32_t *qs, int i)
{
	if (s.gq) return s.q[i];
	const int k = i - s.qbase;
	return (int)(qs[(k >> 4) * kSmemBlock + threadIdx.x] >> ((k & 15) << 1) & 3);
}

__device__ __forceinline__ void set_p(Lane &s, const Sai &p) { s.px0 = p.x0; s.px1 = p.x1; s.psize = p.size; }

// append one record to the task's slots (the spill list past kslot).  PASS 3 / 4 (segment and fix-up
// tasks of segmented pass 1) have no spill: a full task flags its read and is abandoned.  Records of
// every pass-1 kind carry pass 1.
template <int PASS>
__device__ __forceinline__ void emit(const SeedArgs &a, Lane &s)
{
	SeedRec r;
	r.x0 = s.px0; r.x1 = s.px1; r.size = s.psize;
	r.info = (uint64_t)(uint32_t)s.x << 32 | (uint64_t)(uint32_t)s.i;
	r.rid = s.rid; r.pass = PASS == 2 ? 2 : 1;
	if (s.nemit < (uint32_t)a.kslot) {
		a.slots[(uint64_t)s.task * a.kslot + s.nemit] = r;
	} else if (PASS == 3 || PASS == 4) {
		atomicOr(&a.rflag[s.rid], 1u);
		s.stage = kStageAbandon;
		return;
	} else {
		const unsigned long long k = atomicAdd(a.n_spill, 1ULL);
		if ((int64_t)k < a.spill_cap) a.spill[k] = r;
	}
	++s.nemit;
}

// segmented pass 1, fix-up: whether x is one of segment task j's chain values (recorded ascending)
__device__ __forceinline__ bool in_chain(const SeedArgs &a, uint32_t j, uint32_t x)
{
	const uint32_t *c = a.chain + (uint64_t)j * a.chain_cap;
	int lo = 0, hi = (int)a.chain_n[j];
	while (lo < hi) {
		const int mid = (lo + hi) >> 1;
		if (c[mid] < x) lo = mid + 1;
		else hi = mid;
	}
	return lo < (int)a.chain_n[j] && c[lo] == x;
}

// segmented pass 1, fix-up: with x the true chain's value, skip the segments whose lanes joined the
// true chain (x is one of their chain values: their steps from x on are true, and x moves to their
// handoff) and those the true chain passed; false when the read has no segment left
__device__ __forceinline__ bool fix_advance(const SeedArgs &a, Lane &s)
{
	while (s.k < s.K) {
		const uint32_t j = a.segbase[s.rid] + (uint32_t)s.k - 1;
		const int32_t stop = s.k == s.K - 1 ? s.en : (s.k + 1) * a.seg_len;
		if (s.x >= stop) { ++s.k; continue; }                        // conv[j] stays kNoConv
		if (in_chain(a, j, (uint32_t)s.x)) { a.conv[j] = (uint32_t)s.x; s.x = (int32_t)a.hand[j]; ++s.k; continue; }
		return true;
	}
	return false;
}

// the cache key of bases [from-len+1, from]: base `from` in the top two bits, base from-len+1 in the
// lowest (bwt.c:259-261,392-393) -- exactly the little-endian 2-bit field the LDS words hold for those
// bases, so from LDS it is one extraction from at most two words.
__device__ __forceinline__ uint32_t kmer_at(const Lane &s, const uint32_t *qs, int from, int len)
{
	if (s.gq || len > 16) {
		uint32_t z = 0;
		for (int l = 0; l < len; ++l) z = z << 2 | (uint32_t)qget(s, qs, from - l);
		return z;
	}
	const int lo = from - len + 1 - s.qbase;                   // first base, >= 0
	const int w = lo >> 4, sh = (lo & 15) << 1;
	const uint64_t pair = (uint64_t)qs[w * kSmemBlock + threadIdx.x] |
	                      (sh + 2 * len > 32 ? (uint64_t)qs[(w + 1) * kSmemBlock + threadIdx.x] << 32 : 0);
	return (uint32_t)(pair >> sh) & (uint32_t)((1ULL << (2 * len)) - 1);
}

// the end of a segmented pass-1 task's chain: where the next task (or the fix-up) takes over
template <int PASS>
__device__ __forceinline__ void hand_off(const SeedArgs &a, const Lane &s, int32_t x)
{
	if (PASS == 1 && s.stop < s.en) a.hand0[s.rid] = (uint32_t)x;
	if (PASS == 3) a.hand[s.task] = (uint32_t)x, a.chain_n[s.task] = s.nch;
}

// Cheap transitions until the lane needs memory (bwt.c:381-459 minus the rank/lookup steps).
// Segmented pass 1 (PASS 1 for segment 0, 3 for the other segments, 4 for the fix-up) adds, at
// every new start x: the stop at the segment's end, the chain record, and the fix-up's joins.
template <int PASS>
__device__ __forceinline__ int control(const Idx &ix, const SeedArgs &a, Lane &s, const uint32_t *qs)
{
	for (;;) {
		if (s.stage == kStageAbandon) return ACT_DONE;
		if (s.stage == 1) {
			if (PASS == 4 && !fix_advance(a, s)) return ACT_DONE;
			if ((PASS == 1 || PASS == 3) && s.x >= s.stop) { hand_off<PASS>(a, s, s.x); return ACT_DONE; }
			if (PASS == 3) {
				if (s.nch >= (uint32_t)a.chain_cap) { atomicOr(&a.rflag[s.rid], 1u); return ACT_DONE; }
				a.chain[(uint64_t)s.task * a.chain_cap + s.nch++] = (uint32_t)s.x;
			}
			if (s.en - s.x < s.min_len) { hand_off<PASS>(a, s, s.en); return ACT_DONE; }
			if (s.gq) {                                  // only global windows can hold N
				int xn = -1;
				for (int i = s.x; i < s.x + s.min_len; ++i)
					if (s.q[i] > 3) xn = i;
				if (xn >= 0) { s.x = xn + 1; continue; }
			}
			s.i = s.x + s.min_len - 1;
			if (ix.pre_len > 0 && s.min_len >= ix.pre_len) {
				s.kmer = kmer_at(s, qs, s.i, ix.pre_len);
				s.i -= ix.pre_len;
				s.stage = 2;
				return ACT_LOOKUP;
			}
			set_p(s, set_intv(ix, qget(s, qs, s.i)));
			--s.i;
			s.stage = 3;
			continue;
		} else if (s.stage == 3) {
			if (s.i < s.x) { s.i = s.x + s.min_len; s.stage = 4; continue; }
			return ACT_BACK;
		} else if (s.stage == 4) {
			if (s.i == s.en) { emit<PASS>(a, s); hand_off<PASS>(a, s, s.en); return ACT_DONE; }
			if (qget(s, qs, s.i) < 4) return ACT_FWD;
			emit<PASS>(a, s);                           // N: save, skip it, back to stage 1
			if (s.stage == kStageAbandon) return ACT_DONE;
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

// Pack every read into 2-bit words once per batch, one thread per (read, 16-base word), so that a
// lane starting a task copies a few vector loads into LDS instead of re-reading its bases byte by
// byte.  qflag (zeroed before) marks reads the packed form cannot hold: an N, or more than 384 bases.
__global__ void __launch_bounds__(256) pack_reads_kernel(const uint8_t *seq, const int64_t *off, const int32_t *len, int32_t n_reads,
                                                        uint32_t *packed, uint32_t *qflag)
{
	const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t r = g / kPackWords;
	const int w = (int)(g - r * kPackWords);
	if (r >= n_reads) return;
	const int32_t n = len[r];
	if (n > kPackWords * 16) {
		if (w == 0) atomicOr(&qflag[r], 1u);
		return;
	}
	const uint8_t *src = seq + off[r] + w * 16;
	const int m = n - w * 16;
	uint32_t word = 0;
	bool has_n = false;
	#pragma unroll
	for (int j = 0; j < 16; ++j) {
		const uint32_t c = j < m ? src[j] : 0;
		has_n |= c > 3;
		word |= (c & 3) << (j << 1);
	}
	packed[r * kPackWords + w] = word;
	if (has_n) atomicOr(&qflag[r], 1u);
}

// this lane's read into its LDS slots (the whole read, base 0), or the global path for reads with
// an N or longer than the LDS holds
__device__ __forceinline__ void stage_query(const SeedArgs &a, Lane &s, uint32_t *qs, int32_t rlen)
{
	s.qbase = 0;
	s.gq = a.packed == nullptr || a.qflag[s.rid] != 0;
	if (s.gq) return;
	const uint32_t *src = a.packed + (uint64_t)s.rid * kPackWords;
	const int nw = (rlen + 15) >> 4;
	#pragma unroll 1
	for (int w = 0; w < nw; ++w) qs[w * kSmemBlock + threadIdx.x] = src[w];
}

// PASS 1: pass 1 over the reads (task t = read t; segment 0 of a long read); 2: pass 2 over the
// device-built SMEM tasks; 3: segments k >= 1 of long reads; 4: the fix-up of each long read (task t
// = list[t]); 5: the reads whose segment or fix-up tasks overflowed, serially from segment 0's
// handoff, every record to the spill list.
template <int PASS>
__global__ void __launch_bounds__(kSmemBlock) smem_kernel(Idx ix, SeedArgs a)
{
	__shared__ uint32_t qs[kQWords * kSmemBlock];
	Lane s;
	int act = ACT_DONE;
	bool live = false;
	int nblk = 0, nlook = 0, next_ = 0;
	uint32_t n_tasks = a.n_tasks;
	if (a.n_tasks_dev) {
		const int64_t nt = (int64_t)*a.n_tasks_dev;
		n_tasks = (uint32_t)(nt < a.task_cap ? nt : a.task_cap);
	}
	WaveQueue wq;
	for (;;) {
		const uint32_t t = wq.fetch(a.next, act == ACT_DONE);
		if (act == ACT_DONE) {
			if (t >= n_tasks) {
				if (a.stats) {
					atomicAdd(&a.stats[0], (unsigned long long)nblk);
					atomicAdd(&a.stats[1], (unsigned long long)nlook);
					atomicAdd(&a.stats[2], (unsigned long long)next_);
				}
				break;
			}
			int st;
			int32_t rlen;
			s.stop = 0x7fffffff;
			if (PASS == 1) {
				s.rid = t;
				s.q = a.seq + a.off[t];
				st = 0; s.en = rlen = a.len[t];
				s.min_len = a.min_len; s.min_occ = 1;
				if (a.seg_len > 0 && rlen >= 2 * a.seg_len) s.stop = a.seg_len;
			} else if (PASS == 3) {
				s.rid = a.seg_rid[t];
				s.q = a.seq + a.off[s.rid];
				s.en = rlen = a.len[s.rid];
				s.min_len = a.min_len; s.min_occ = 1;
				s.K = rlen / a.seg_len;
				s.k = (int32_t)(t - a.segbase[s.rid]) + 1;
				st = s.k * a.seg_len;
				s.stop = s.k == s.K - 1 ? rlen : (s.k + 1) * a.seg_len;
				s.nch = 0;
				a.conv[t] = kNoConv;
			} else if (PASS == 4 || PASS == 5) {
				s.rid = a.list[t];
				s.q = a.seq + a.off[s.rid];
				s.en = rlen = a.len[s.rid];
				s.min_len = a.min_len; s.min_occ = 1;
				s.K = rlen / a.seg_len;
				s.k = 1;
				st = (int32_t)a.hand0[s.rid];
			} else {
				const SeedTask tk = a.tasks[t];
				s.rid = tk.rid;
				s.q = a.seq + a.off[tk.rid];
				rlen = a.len[tk.rid];
				st = tk.st; s.en = tk.en;
				s.min_len = tk.min_len; s.min_occ = (uint64_t)(int64_t)tk.min_occ;
			}
			s.task = t;
			s.nemit = 0;
			s.x = st;
			s.stage = PASS == 4 && a.rflag[s.rid] ? kStageAbandon : 1;
			live = true;
			stage_query(a, s, qs, rlen);
		}
		act = control<PASS>(ix, a, s, qs);
		if (act == ACT_DONE) {
			if (live && a.slot_cnt) a.slot_cnt[s.task] = s.nemit < (uint32_t)a.kslot ? s.nemit : (uint32_t)a.kslot;
			live = false;
			continue;
		}
		if (act == ACT_LOOKUP) {                        // stages 2/5 (bwt.c:401-413)
			const Sai p = ix.pre[s.kmer];
			++nlook;
			if (p.size < s.min_occ) {
				s.i += ix.pre_len;
				set_p(s, set_intv(ix, qget(s, qs, s.i)));
				--s.i;
			} else set_p(s, p);
			++s.stage;
			continue;
		}
		// one extension per trip: backward for stages 3/6 (se_one_step_back, bwt.c:339-354),
		// forward for stage 4 (bwt.c:421-453)
		const int back = act == ACT_BACK;
		const int qc = qget(s, qs, s.i);
		Sai cur;
		cur.x0 = s.px0; cur.x1 = s.px1; cur.size = s.psize; cur.info = 0;
		const Sai o = extend1(ix, cur, back, back ? qc : 3 - qc, &nblk);
		++next_;
		if (back) {
			if (o.size < s.min_occ) { s.x = s.i + 1; s.stage = 1; }
			else { set_p(s, o); --s.i; }
		} else if (o.size >= s.min_occ) {
			set_p(s, o); ++s.i;
		} else {
			emit<PASS>(a, s);
			if (s.stage == kStageAbandon) continue;
			if (ix.pre_len > 0 && s.i - s.x - 1 >= ix.pre_len) {
				s.kmer = kmer_at(s, qs, s.i, ix.pre_len);
				s.i -= ix.pre_len;
				s.stage = 5;
			} else {
				set_p(s, set_intv(ix, qget(s, qs, s.i)));
				--s.i;
				s.stage = 6;
			}
		}
	}
}

// pass-2 tasks (seed.c:80-98): every pass-1 SMEM with length >= 2*min_len and size <= max_sub_occ.
__device__ __forceinline__ bool subtask_ok(const SeedRec &r, int32_t min_len, int32_t max_sub_occ)
{
	const uint32_t st = (uint32_t)(r.info >> 32), en = (uint32_t)r.info;
	return !(en - st < (uint32_t)(min_len * 2) || r.size > (uint64_t)(int64_t)max_sub_occ);
}

__device__ __forceinline__ SeedTask make_subtask(const SeedRec &r, int32_t min_len)
{
	const uint32_t st = (uint32_t)(r.info >> 32), en = (uint32_t)r.info;
	SeedTask t;
	t.rid = r.rid;
	t.st = (int32_t)st; t.en = (int32_t)en;
	t.min_len = (int32_t)(en - st) / 2 > min_len ? (int32_t)(en - st) / 2 : min_len;
	t.min_occ = (int32_t)r.size + 1;
	return t;
}

// one wave per 64 reads: over the kslot steps, lane l takes slot (step*64 + l) % kslot of read
// (step*64 + l) / kslot, so the wave reads its reads' slot arrays contiguously.  A counting pass and
// one atomic give the wave a contiguous range of task positions in (read, slot) order: the tasks of
// neighbouring reads stay together, which pass 2's FM-index accesses profit from.
__device__ __forceinline__ bool subtask_at(const SeedRec *slots, const uint32_t *slot_cnt, int32_t kslot, int32_t n_reads, int64_t r0,
                                           int64_t g, int32_t min_len, int32_t max_sub_occ, SeedTask &t)
{
	const int64_t rd = r0 + g / kslot;
	const int32_t sj = (int32_t)(g % kslot);
	if (rd >= n_reads || sj >= (int32_t)slot_cnt[rd]) return false;
	const SeedRec r = slots[rd * kslot + sj];
	if (!subtask_ok(r, min_len, max_sub_occ)) return false;
	t = make_subtask(r, min_len);
	return true;
}

__global__ void __launch_bounds__(256) subtask_kernel(const SeedRec *slots, const uint32_t *slot_cnt, int32_t kslot, int32_t n_reads,
                                                     const SeedRec *spill, const unsigned long long *n_spill, int64_t spill_cap,
                                                     int32_t min_len, int32_t max_sub_occ, SeedTask *tasks, uint32_t *n_tasks, int64_t task_cap)
{
	const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int lane = threadIdx.x & 63;
	const int64_t r0 = (tid >> 6) << 6;                   // first read of this wave
	uint32_t tot = 0;
	for (int64_t g = lane; g < (int64_t)64 * kslot; g += 64) {
		SeedTask t;
		tot += (uint32_t)__popcll(__ballot(subtask_at(slots, slot_cnt, kslot, n_reads, r0, g, min_len, max_sub_occ, t)));
	}
	uint32_t base = 0;
	if (lane == 0 && tot) base = atomicAdd(n_tasks, tot);
	base = (uint32_t)__shfl((int)base, 0);
	for (int64_t g = lane; tot && g < (int64_t)64 * kslot; g += 64) {
		SeedTask t;
		const bool q = subtask_at(slots, slot_cnt, kslot, n_reads, r0, g, min_len, max_sub_occ, t);
		const uint64_t m = __ballot(q);
		const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
		if (q && (int64_t)pos < task_cap) tasks[pos] = t;
		base += (uint32_t)__popcll(m);
	}
	int64_t ns = (int64_t)*n_spill;
	if (ns > spill_cap) ns = spill_cap;
	for (int64_t k = tid; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
		const SeedRec r = spill[k];
		if (r.pass != 1 || !subtask_ok(r, min_len, max_sub_occ)) continue;
		const uint32_t t = atomicAdd(n_tasks, 1u);
		if ((int64_t)t < task_cap) tasks[t] = make_subtask(r, min_len);
	}
}

// records of the two slot arrays (cnt/off over n1 + n2 tasks, exclusive offsets) and the spill list
// into one flat array; *n_out = total
__global__ void __launch_bounds__(256) compact_kernel(const SeedRec *slots1, const SeedRec *slots2, int32_t k1, int32_t k2, int64_t n1,
                                                     int64_t n_cnt, const uint32_t *cnt, const uint32_t *off, const SeedRec *spill,
                                                     const unsigned long long *n_spill, int64_t spill_cap, SeedRec *out, int64_t cap,
                                                     int64_t *n_out)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t total = n_cnt > 0 ? (uint64_t)off[n_cnt - 1] + cnt[n_cnt - 1] : 0;
	if (i < n_cnt) {
		const uint32_t c = cnt[i], o = off[i];
		const SeedRec *src = i < n1 ? slots1 + (uint64_t)i * k1 : slots2 + (uint64_t)(i - n1) * k2;
		for (uint32_t j = 0; j < c; ++j)
			if ((int64_t)(o + j) < cap) out[o + j] = src[j];
	}
	int64_t ns = (int64_t)*n_spill;
	if (ns > spill_cap) ns = spill_cap;
	for (int64_t k = i; k < ns; k += (int64_t)gridDim.x * blockDim.x)
		if ((int64_t)(total + k) < cap) out[total + k] = spill[k];
	if (i == 0) *n_out = (int64_t)(total + (uint64_t)ns);
}

// ---------------------------------------------------------------------------------------------
// SA locate (mb_bwt_sa / mb_bwt_sa_batch, bwt.c:499-541): LF-walk each row to the nearest sampled
// row; lanes refill from a counter because walk lengths are geometric (mean 2^sa_bit).
// ---------------------------------------------------------------------------------------------
// SA of rows i << fine_bit: the walk of mb_bwt_sa_batch (bwt.c:511-541) from each, so that a walk
// from any row may stop at the first row of that finer grid: the result is steps + fine value, the
// same number the walk to a sampled row gives
__global__ void __launch_bounds__(256) sa_fine_kernel(Idx ix, uint64_t *fine, uint64_t n_fine)
{
	const uint64_t mask = (1ULL << ix.sa_bit) - 1;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_fine; i += (uint64_t)gridDim.x * blockDim.x) {
		uint64_t k = i << ix.fine_bit, steps = 0;
		while ((k & mask) != 0) k = lf(ix, k), ++steps;
		fine[i] = steps + ix.sa[k >> ix.sa_bit];
	}
}

__global__ void __launch_bounds__(256) sa_kernel(Idx ix, int64_t n, const int64_t *n_dev, uint64_t *x, uint32_t *next,
                                                unsigned long long *stats)
{
	const uint32_t bit = ix.sa_fine ? ix.fine_bit : ix.sa_bit;
	const uint64_t *samp = ix.sa_fine ? ix.sa_fine : ix.sa;
	const uint64_t mask = (1ULL << bit) - 1;
	if (n_dev) { const int64_t nd = *n_dev; n = nd < n ? nd : n; }
	bool have = false;
	uint32_t idx = 0;
	uint64_t k = 0, steps = 0, tot = 0;
	WaveQueue wq;
	for (;;) {
		const uint32_t t = wq.fetch(next, !have);
		if (!have) {
			if ((int64_t)t >= n) break;
			idx = t; k = x[t]; steps = 0; have = true;
			if (k > ix.seq_len) {          // not a row of the index: no walk, the result says so
				x[idx] = ~0ULL;
				have = false;
				continue;
			}
		}
		if ((k & mask) == 0) {
			x[idx] = steps + samp[k >> bit];
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
	if ((uint64_t)en > clen) en = (int64_t)clen;   // unsigned, as l2bit.c:61 compares: a negative en clips to the end
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

// dependent random 64-byte block fetches over the occ table (random-access ceiling)
__global__ void gather_chain_kernel(Idx ix, int64_t n_chains, int32_t steps, unsigned long long *sum)
{
	const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (id >= n_chains) return;
	const uint64_t nblk = (ix.seq_len + 127) >> 7;
	uint64_t x = ((uint64_t)id * 0x9e3779b97f4a7c15ULL) % nblk, acc = 0;
	for (int s = 0; s < steps; ++s) {
		uint64_t b[8];
		load_block(ix.data, x, b);
		acc += b[0] ^ b[5];
		x = ((x ^ (b[1] + b[6])) * 0xbf58476d1ce4e5b9ULL + (uint64_t)s) % nblk;
	}
	atomicAdd(sum, (unsigned long long)acc);
}

hipError_t launch_gather_chain(const Idx &ix, int64_t n_chains, int32_t steps, int32_t bs, unsigned long long *sum, hipStream_t st)
{
	hipLaunchKernelGGL(gather_chain_kernel, dim3((unsigned)((n_chains + bs - 1) / bs)), dim3(bs), 0, st, ix, n_chains, steps, sum);
	return hipGetLastError();
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

hipError_t launch_pack_reads(const uint8_t *seq, const int64_t *off, const int32_t *len, int32_t n_reads, uint32_t *packed,
                             uint32_t *qflag, hipStream_t st)
{
	const int64_t n = (int64_t)n_reads * kPackWords;
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(pack_reads_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, seq, off, len, n_reads, packed, qflag);
	return hipGetLastError();
}

hipError_t launch_kmer_cache(const Idx &ix, Sai *pre, int len, hipStream_t st)
{
	const uint64_t n = 1ULL << (2 * len);
	// a dispatch holds fewer than 2^32 work-items: larger tables are covered by a grid-stride loop
	const int64_t nb = blocks_for((int64_t)n, 256);
	hipLaunchKernelGGL(kmer_cache_kernel, dim3((unsigned)(nb < (1 << 22) ? nb : (1 << 22))), dim3(256), 0, st, ix, pre, len, n);
	return hipGetLastError();
}

hipError_t launch_sa_fine(const Idx &ix, uint64_t *fine, uint64_t n_fine, hipStream_t st)
{
	const int64_t nb = blocks_for((int64_t)n_fine, 256);
	hipLaunchKernelGGL(sa_fine_kernel, dim3((unsigned)(nb < (1 << 22) ? nb : (1 << 22))), dim3(256), 0, st, ix, fine, n_fine);
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
	switch (pass) {
	case 1: hipLaunchKernelGGL(smem_kernel<1>, dim3(grid), dim3(256), 0, st, ix, a); break;
	case 2: hipLaunchKernelGGL(smem_kernel<2>, dim3(grid), dim3(256), 0, st, ix, a); break;
	case 3: hipLaunchKernelGGL(smem_kernel<3>, dim3(grid), dim3(256), 0, st, ix, a); break;
	case 4: hipLaunchKernelGGL(smem_kernel<4>, dim3(grid), dim3(256), 0, st, ix, a); break;
	case 5: hipLaunchKernelGGL(smem_kernel<5>, dim3(grid), dim3(256), 0, st, ix, a); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

// ---- segmented pass 1: task lists and the merge ----------------------------------------------

__global__ void __launch_bounds__(256) seg_count_kernel(const int32_t *len, int32_t n_reads, int32_t seg_len, uint32_t *nseg)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= n_reads) return;
	const int32_t l = len[r];
	nseg[r] = l >= 2 * seg_len ? (uint32_t)(l / seg_len - 1) : 0u;
}

// per long read: its segment tasks' read ids at segbase[r]..; the read into the long-read list (room
// for every read).  A read whose segment tasks, or whose fix-up task, fall past seg_cap is flagged:
// it gets the serial fallback, so a full pool costs time, never results.
__global__ void __launch_bounds__(256) seg_build_kernel(const int32_t *len, int32_t n_reads, int32_t seg_len, const uint32_t *segbase,
                                                       int64_t seg_cap, uint32_t *seg_rid, uint32_t *longl, uint32_t *n_long,
                                                       uint32_t *rflag)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= n_reads) return;
	const int32_t l = len[r];
	if (l < 2 * seg_len) return;
	const uint32_t b = segbase[r], n = (uint32_t)(l / seg_len - 1);
	for (uint32_t k = 0; k < n; ++k)
		if ((int64_t)(b + k) < seg_cap) seg_rid[b + k] = (uint32_t)r;
	const uint32_t i = atomicAdd(n_long, 1u);
	longl[i] = (uint32_t)r;
	if ((int64_t)(b + n) > seg_cap || (int64_t)i >= seg_cap) rflag[r] = 1u;
}

__global__ void __launch_bounds__(256) seg_flagged_kernel(const uint32_t *longl, const uint32_t *n_long, const uint32_t *rflag,
                                                         uint32_t *fbl, uint32_t *n_fb)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= *n_long) return;
	const uint32_t r = longl[i];
	if (rflag[r]) fbl[atomicAdd(n_fb, 1u)] = r;
}

// one wave per long read: segment k's records from where its lane joined the true chain, and the
// fix-up's records, appended to the spill list (a flagged read's come from its fallback task instead)
__global__ void __launch_bounds__(64) seg_merge_kernel(const int32_t *len, int32_t seg_len, const uint32_t *longl, const uint32_t *n_long,
                                                      const uint32_t *rflag, const uint32_t *segbase, const SeedRec *segslots,
                                                      const uint32_t *segcnt, int32_t kseg, const uint32_t *conv, const SeedRec *fixslots,
                                                      const uint32_t *fixcnt, int32_t kfix, SeedRec *spill, unsigned long long *n_spill,
                                                      int64_t spill_cap)
{
	const uint32_t i = blockIdx.x;                 // the grid is the fix-up cap: later entries are flagged
	if (i >= *n_long) return;
	const uint32_t r = longl[i];
	if (rflag[r]) return;
	const uint32_t b = segbase[r], n = (uint32_t)(len[r] / seg_len - 1);
	for (uint32_t j = b; j < b + n; ++j)
		for (uint32_t e = threadIdx.x; e < segcnt[j]; e += 64) {
			const SeedRec rec = segslots[(uint64_t)j * kseg + e];
			if ((uint32_t)(rec.info >> 32) < conv[j]) continue;
			const unsigned long long k = atomicAdd(n_spill, 1ULL);
			if ((int64_t)k < spill_cap) spill[k] = rec;
		}
	for (uint32_t e = threadIdx.x; e < fixcnt[i]; e += 64) {
		const unsigned long long k = atomicAdd(n_spill, 1ULL);
		if ((int64_t)k < spill_cap) spill[k] = fixslots[(uint64_t)i * kfix + e];
	}
}

hipError_t launch_seg_count(const int32_t *len, int32_t n_reads, int32_t seg_len, uint32_t *nseg, hipStream_t st)
{
	if (n_reads <= 0) return hipSuccess;
	hipLaunchKernelGGL(seg_count_kernel, dim3(blocks_for(n_reads, 256)), dim3(256), 0, st, len, n_reads, seg_len, nseg);
	return hipGetLastError();
}

hipError_t launch_seg_build(const int32_t *len, int32_t n_reads, int32_t seg_len, const uint32_t *segbase, int64_t seg_cap,
                            uint32_t *seg_rid, uint32_t *longl, uint32_t *n_long, uint32_t *rflag, hipStream_t st)
{
	if (n_reads <= 0) return hipSuccess;
	hipLaunchKernelGGL(seg_build_kernel, dim3(blocks_for(n_reads, 256)), dim3(256), 0, st, len, n_reads, seg_len, segbase, seg_cap,
	                   seg_rid, longl, n_long, rflag);
	return hipGetLastError();
}

hipError_t launch_seg_flagged(const uint32_t *longl, const uint32_t *n_long, const uint32_t *rflag, uint32_t *fbl, uint32_t *n_fb,
                              int64_t cap, hipStream_t st)
{
	hipLaunchKernelGGL(seg_flagged_kernel, dim3(blocks_for(cap, 256)), dim3(256), 0, st, longl, n_long, rflag, fbl, n_fb);
	return hipGetLastError();
}

hipError_t launch_seg_merge(const int32_t *len, int32_t seg_len, const uint32_t *longl, const uint32_t *n_long, const uint32_t *rflag,
                            const uint32_t *segbase, const SeedRec *segslots, const uint32_t *segcnt, int32_t kseg,
                            const uint32_t *conv, const SeedRec *fixslots, const uint32_t *fixcnt, int32_t kfix, SeedRec *spill,
                            unsigned long long *n_spill, int64_t spill_cap, int64_t cap, hipStream_t st)
{
	hipLaunchKernelGGL(seg_merge_kernel, dim3((unsigned)cap), dim3(64), 0, st, len, seg_len, longl, n_long, rflag, segbase, segslots,
	                   segcnt, kseg, conv, fixslots, fixcnt, kfix, spill, n_spill, spill_cap);
	return hipGetLastError();
}

hipError_t launch_subtask(const SeedRec *slots, const uint32_t *slot_cnt, int32_t kslot, int32_t n_reads, const SeedRec *spill,
                          const unsigned long long *n_spill, int64_t spill_cap, int32_t min_len, int32_t max_sub_occ,
                          SeedTask *tasks, uint32_t *n_tasks, int64_t task_cap, hipStream_t st)
{
	const unsigned grid = (unsigned)((n_reads + 255) / 256) + 1;
	hipLaunchKernelGGL(subtask_kernel, dim3(grid), dim3(256), 0, st, slots, slot_cnt, kslot, n_reads, spill, n_spill, spill_cap,
	                   min_len, max_sub_occ, tasks, n_tasks, task_cap);
	return hipGetLastError();
}

hipError_t launch_compact(const SeedRec *slots1, const SeedRec *slots2, int32_t k1, int32_t k2, int64_t n1, int64_t n_cnt,
                          const uint32_t *cnt, const uint32_t *off, const SeedRec *spill, const unsigned long long *n_spill,
                          int64_t spill_cap, SeedRec *out, int64_t cap, int64_t *n_out, hipStream_t st)
{
	const unsigned grid = (unsigned)((n_cnt + 255) / 256) + 1;
	hipLaunchKernelGGL(compact_kernel, dim3(grid), dim3(256), 0, st, slots1, slots2, k1, k2, n1, n_cnt, cnt, off, spill, n_spill,
	                   spill_cap, out, cap, n_out);
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
