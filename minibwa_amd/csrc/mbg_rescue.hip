// mbg_rescue.hip -- mate rescue on gfx950: for each candidate window the reference's
// mb_matesw_core step (pe.c:342-349), mb_ungap's k-mer vote (pe.c:209-247), the window narrowing,
// and the job building around mb_matesw_align (pe.c:249-312): ksw_ll runs in mbg_ll.hip and the
// extension ksw_extz2 of the reversed prefixes in mbg_dp.hip; this file builds their jobs and
// assembles the hit fields.
//
// mb_ungap on a wavefront: the query's 7-mers go into an open-addressing table in LDS keyed by the
// k-mer and holding the LAST position where it ends (an LDS atomicMax on key<<16|pos, so the order
// in which lanes insert does not matter); every window k-mer then looks itself up and adds one
// vote to its diagonal, two 16-bit counters a dword in LDS.  The top vote, its first diagonal and
// the count of diagonals above half of it are order-free reductions.
#include "mbg_internal.h"
#include "mbg_wave.h"

namespace mbg {

// base conversion of methylation mode mt as four bytes (0: none, 1: C->T, 2: G->A; pe.c:211-212)
__device__ __forceinline__ uint32_t ug_conv(int mt) { return mt == 1 ? 0x03020300u : mt == 2 ? 0x03000100u : 0x03020100u; }

// the 7-mer ending at s[i] (oldest base in the high bits), false if it holds an N or starts before s
__device__ __forceinline__ bool kmer7_at(const uint8_t *s, int i, uint32_t conv, uint32_t *x)
{
	if (i < 6) return false;
	uint32_t v = 0;
#pragma unroll
	for (int j = 0; j < 7; ++j) {
		const uint32_t c = s[i - 6 + j];
		if (c > 3) return false;
		v = v << 2 | (conv >> (c << 3) & 3);
	}
	*x = v;
	return true;
}

__device__ __forceinline__ uint32_t ug_slot(uint32_t x, int T) { return (x * 0x9E3779B1u >> 16) & (uint32_t)(T - 1); }

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
	for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// one wavefront per candidate; LDS: T table slots, then ceil(max window / 2) counter dwords
__global__ __launch_bounds__(64) void ungap_kernel(MsArgs a)
{
	extern __shared__ uint32_t ug_lds[];
	const int c = blockIdx.x, lane = threadIdx.x;
	if (c >= a.n) return;
	const MsJob jb = a.jobs[c];
	const int qlen = jb.qlen, wlen = (int)a.wlen[c];
	const uint8_t *q = a.qseq + jb.qoff, *w = a.win + a.woff[c];
	const int T = a.tab_size;
	const uint32_t conv = ug_conv(a.mt);
	uint32_t *tab = ug_lds, *cnt = ug_lds + T;
	MsRes *r = a.res + c;
	LlJob lj{jb.qoff, a.woff[c], -1, 0, 1, 0};              // qlen -1: ksw_ll skips the job
	if (wlen <= qlen) {                                      // not a candidate (pe.c:342)
		if (lane == 0) {
			r->max_ug = 0, r->max_i = -1, r->n_good = 0, r->n_kmer = 0, r->ts2 = 0, r->te2 = 0, r->aligned = 0;
			a.ll_jobs[c] = lj;
		}
		return;
	}
	for (int i = lane; i < T; i += 64) tab[i] = 0;
	for (int i = lane; i < (wlen + 1) >> 1; i += 64) cnt[i] = 0;
	__syncthreads();
	int fresh = 0;
	for (int i = lane; i < qlen; i += 64) {
		uint32_t x;
		if (!kmer7_at(q, i, conv, &x)) continue;
		const uint32_t key = (x + 1) << 16, v = key | (uint32_t)i;
		uint32_t s = ug_slot(x, T);
		for (;;) {                                           // T >= 2*qlen: a free slot always exists
			uint32_t cur = tab[s];
			if (cur == 0) {
				cur = atomicCAS(&tab[s], 0u, v);
				if (cur == 0) { ++fresh; break; }
			}
			if ((cur & 0xffff0000u) == key) { atomicMax(&tab[s], v); break; }
			s = (s + 1) & (uint32_t)(T - 1);
		}
	}
	const int n_kmer = wave_sum(fresh);
	__syncthreads();
	for (int j = lane; j < wlen; j += 64) {
		uint32_t x, cur;
		if (!kmer7_at(w, j, conv, &x)) continue;
		const uint32_t key = (x + 1) << 16;
		uint32_t s = ug_slot(x, T);
		while ((cur = tab[s]) != 0 && (cur & 0xffff0000u) != key) s = (s + 1) & (uint32_t)(T - 1);
		if (cur == 0) continue;
		const int d = j - (int)(cur & 0xffff);
		if (d >= 0) atomicAdd(&cnt[d >> 1], 1u << ((d & 1) << 4));
	}
	__syncthreads();
	int best = 0, first = 0x7fffffff;
	for (int d = lane; d < wlen; d += 64) {
		const int v = (int)(cnt[d >> 1] >> ((d & 1) << 4) & 0xffff);
		if (v > best) best = v, first = d;
	}
	const int top = (int)wave_max_u32((uint32_t)best);
	const int max_i = top > 0 ? (int)~wave_max_u32(best == top ? ~(uint32_t)first : 0u) : -1;
	int good = 0;
	for (int d = lane; d < wlen; d += 64) good += (int)(cnt[d >> 1] >> ((d & 1) << 4) & 0xffff) > top >> 1;
	const int n_good = wave_sum(good);
	if (lane != 0) return;
	// pe.c:343-349: narrow around a clear diagonal, align when the vote passes
	int ts2 = 0, te2 = wlen;
	if (top >= 10 && top >= qlen >> 1 && n_good == 1) {
		ts2 = max_i - qlen / 2, te2 = ts2 + qlen * 2;
		if (ts2 < 0) ts2 = 0;
		if (te2 > wlen) te2 = wlen;
	}
	const bool aligned = top >= 10 || (double)top >= n_kmer * 0.33;
	r->max_ug = top, r->max_i = max_i, r->n_good = n_good, r->n_kmer = n_kmer, r->ts2 = ts2, r->te2 = te2, r->aligned = aligned;
	// mb_matesw_align's ksw_ll (pe.c:254-270): u8 core while the best score stays below 255 - mismatch
	const int max_sc = qlen < te2 - ts2 ? qlen : te2 - ts2;
	if (aligned && max_sc < 32767) {
		lj.toff = a.woff[c] + ts2, lj.qlen = qlen, lj.tlen = te2 - ts2, lj.size = max_sc < 255 - a.b_mm ? 1 : 2;
	}
	a.ll_jobs[c] = lj;
}

// one wavefront per candidate: the extension job on the reversed prefixes ending at the ksw_ll ends
// (pe.c:274-287), reversed copies written into rq/rt at the candidate's own offsets
__global__ __launch_bounds__(64) void rescue_dp_kernel(MsArgs a)
{
	const int c = blockIdx.x, lane = threadIdx.x;
	if (c >= a.n) return;
	const MsJob jb = a.jobs[c];
	const LlJob lj = a.ll_jobs[c];
	const LlRes ll = a.ll_res[c];
	const bool ran = lj.qlen >= 0 && ll.score >= a.min_dp_max && ll.score >= jb.min_sc;
	DpJob d{};
	d.qoff = jb.qoff, d.toff = lj.toff, d.cigar_off = jb.cigar_off;
	d.q = d.q2 = (int8_t)a.q, d.e = d.e2 = (int8_t)a.e;
	d.w = a.bw, d.zdrop = a.zdrop, d.end_bonus = a.end_bonus;
	d.flag = 0x01;                                            // score only: an empty job
	if (ran) {
		const int qe = ll.qe + 1, te = ll.te + 1;
		const uint8_t *q = a.qseq + jb.qoff, *t = a.win + lj.toff;
		for (int i = lane; i < qe; i += 64) a.rq[jb.qoff + i] = q[qe - 1 - i];
		for (int i = lane; i < te; i += 64) a.rt[lj.toff + i] = t[te - 1 - i];
		d.qlen = qe, d.tlen = te;
		d.flag = 0x40 | 0x02 | 0x80 | (a.mt ? 0x04 : 0);      // EXTZ_ONLY|RIGHT|REV_CIGAR (+GENERIC_SC)
	}
	if (lane == 0) {
		a.dp_jobs[c] = d;
		a.res[c].ran_dp = ran;
	}
}

// one thread per candidate: the hit fields before mb_update_extra (pe.c:288-300), and the number of
// CIGAR operations to bring back
__global__ void rescue_finish_kernel(MsArgs a)
{
	// dp_max2's product and sum rounded one at a time, as the host code computes them
#pragma clang fp contract(off)
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= a.n) return;
	MsRes &o = a.res[c];
	const LlJob lj = a.ll_jobs[c];
	const LlRes ll = lj.qlen >= 0 ? a.ll_res[c] : LlRes{-1, -1, -1, -1, -1};
	const DpRes ez = o.ran_dp ? a.dp_res[c] : DpRes{};
	o.ll = ll, o.ez = ez;
	a.ccnt[c] = o.ran_dp && ez.n_cigar > 0 ? (uint32_t)ez.n_cigar : 0u;
	o.rescued = o.ran_dp && ez.n_cigar > 0 && ez.max >= a.min_dp_max * a.a;
	o.qs = o.qe = o.ts = o.te = o.dp_max2 = 0;
	o.n_cigar = o.blen = o.mlen = o.n_ambi = o.dp_max = 0;
	ExtraJob x{};
	x.cigar_off = a.jobs[c].cigar_off;
	if (o.rescued) {
		const int qe = ll.qe + 1, te = ll.te + 1;
		o.qe = qe, o.te = te;
		o.ts = te - (ez.reach_end ? ez.mqe_t + 1 : ez.max_t + 1);
		o.qs = qe - (ez.reach_end ? qe : ez.max_q + 1);
		const int m2 = (int)((double)ez.max / ll.score * ll.score2 + .499);
		o.dp_max2 = m2 < 0 ? 0 : m2;
		// mb_update_extra on the hit (pe.c:309): the read from qs, the aligned window from ts
		x.qoff = a.jobs[c].qoff + o.qs, x.toff = lj.toff + o.ts;
		x.qs = o.qs, x.qe = o.qe, x.ts = o.ts, x.te = o.te, x.rev = 0, x.n_cigar = ez.n_cigar;
	}
	a.extra[c] = x;
}

// one thread per candidate: the hit fields mb_update_extra left, and the CIGAR length to bring back
__global__ void rescue_final_kernel(MsArgs a)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= a.n) return;
	MsRes &o = a.res[c];
	if (!o.rescued) return;
	const ExtraJob x = a.extra[c];
	o.qs = x.qs, o.qe = x.qe, o.ts = (int32_t)x.ts, o.te = (int32_t)x.te;
	o.n_cigar = x.n_cigar, o.blen = x.blen, o.mlen = x.mlen, o.n_ambi = x.n_ambi, o.dp_max = x.dp_max;
	a.ccnt[c] = (uint32_t)x.n_cigar;
}

hipError_t launch_rescue_final(const MsArgs &a, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	hipLaunchKernelGGL(rescue_final_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
	return hipGetLastError();
}

// one thread per candidate: its CIGAR moved from the DP's slots to a packed list at ccnt's prefix sum
__global__ void rescue_cigar_kernel(MsArgs a, const uint32_t *coff, uint32_t *packed)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= a.n) return;
	const uint32_t n = a.ccnt[c];
	const uint32_t *src = a.cigar + a.jobs[c].cigar_off;
	for (uint32_t i = 0; i < n; ++i) packed[coff[c] + i] = src[i];
}

hipError_t launch_rescue_cigar(const MsArgs &a, const uint32_t *coff, uint32_t *packed, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	hipLaunchKernelGGL(rescue_cigar_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a, coff, packed);
	return hipGetLastError();
}

hipError_t launch_ungap(const MsArgs &a, int max_wlen, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	const size_t shm = (size_t)a.tab_size * 4 + (size_t)((max_wlen + 1) / 2) * 4;
	hipLaunchKernelGGL(ungap_kernel, dim3((unsigned)a.n), dim3(64), shm, st, a);
	return hipGetLastError();
}

hipError_t launch_rescue_dp(const MsArgs &a, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	hipLaunchKernelGGL(rescue_dp_kernel, dim3((unsigned)a.n), dim3(64), 0, st, a);
	return hipGetLastError();
}

hipError_t launch_rescue_finish(const MsArgs &a, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	hipLaunchKernelGGL(rescue_finish_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
	return hipGetLastError();
}

} // namespace mbg
