// mbg_align.hip -- mb_align_pair (align.c:316-390) over a batch: the decisions made before the DP.
//
// One wave per call.  The reference scans the pair base by base for its two ungapped fast paths;
// here the 64 lanes take 64 positions at a time: the running score is a wave prefix sum, the
// extension's maximum and its first position a wave max plus a ballot, and the mismatch count a
// ballot popcount.  A call the fast paths settle gets its ksw_extz_t fields and its one-operation
// CIGAR here; a call the max_sw_mat rule skips gets the reset fields with zdropped set; every other
// call becomes a DP job (band capped, z-drop scaled, generic-scoring bit added) for the batch's
// extz2 or extd2 kernels, and align_merge_kernel puts the decided results over the DP's empty
// results for the jobs it skipped.
#include "mbg_internal.h"
#include "mbg_wave.h"

namespace mbg {

// inclusive prefix sum over the 64 lanes (the pattern of wave_scan_max_u32 with adds; lanes a DPP
// step does not reach add 0)
__device__ __forceinline__ int32_t wave_scan_add_i32(int32_t v)
{
	v += (int32_t)dpp_u32<0x111>((uint32_t)v);
	v += (int32_t)dpp_u32<0x112>((uint32_t)v);
	v += (int32_t)dpp_u32<0x114>((uint32_t)v);
	v += (int32_t)dpp_u32<0x118>((uint32_t)v);
	v += (int32_t)dpp_u32<0x142, 0xa>((uint32_t)v);
	v += (int32_t)dpp_u32<0x143, 0xc>((uint32_t)v);
	return v;
}

constexpr int32_t kAlnNegInf = -0x40000000;   // KSW_NEG_INF (ksw2.h:8)

// ksw_reset_extz (ksw2.h:174-179)
__device__ __forceinline__ void reset_res(DpRes &r)
{
	r.max_q = r.max_t = r.mqe_t = r.mte_q = -1;
	r.max = 0, r.score = r.mqe = r.mte = kAlnNegInf;
	r.n_cigar = 0, r.zdropped = 0, r.reach_end = 0, r.cells = 0;
}

__global__ void __launch_bounds__(256) align_plan_kernel(AlignArgs a)
{
	__shared__ int8_t mat[32];
	if (threadIdx.x < 25) mat[threadIdx.x] = a.mat[threadIdx.x];
	__syncthreads();
	const int lane = threadIdx.x & 63;
	const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
	for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < a.n; i += nw) {
		const AlignJob jb = a.jobs[i];
		const int qlen = jb.qlen, tlen = jb.tlen;
		int flag = jb.flag;
		if (a.generic) flag |= 0x04;                                   // align.c:334-335
		const uint8_t *qs = a.qseq + jb.qoff, *ts = a.tseq + jb.toff;
		const bool ext = (flag & 0x40) && tlen >= qlen;                 // align.c:336
		const bool glob = !ext && qlen == tlen && !(flag & 0x40);      // align.c:351
		DpRes r;
		reset_res(r);
		int kind = 0;                                                  // 0 DP, 1 settled here
		int n_mm = -1;
		if (ext || glob) {
			int32_t base = 0, best = 0, best_j = -1, mm = 0;
			for (int j0 = 0; j0 < qlen; j0 += 64) {
				const int j = j0 + lane;
				int32_t sc = 0;
				bool bad = false;
				if (j < qlen) {
					const uint8_t tj = ts[j], qj = qs[j];
					sc = mat[(tj < 4 ? tj : 4) * 5 + (qj < 4 ? qj : 4)];   // codes above 4 score as N
					bad = tj > 3 || qj > 3 || sc < 0;
				}
				mm += __popcll(__ballot(bad));
				const int32_t S = base + wave_scan_add_i32(sc);   // score after position j
				if (ext) {                                       // the first position of a new maximum
					const uint32_t key = j < qlen ? (uint32_t)S ^ 0x80000000u : 0u;
					const int32_t cm = (int32_t)(wave_max_u32(key) ^ 0x80000000u);
					if (cm > best) {
						const uint64_t at = __ballot(j < qlen && S == cm);
						best = cm, best_j = j0 + (int)__ffsll((unsigned long long)at) - 1;
					}
				}
				base = __builtin_amdgcn_readlane(S, 63);
			}
			n_mm += mm;
			r.score = base;
			if (ext) {
				r.max = best, r.max_q = r.max_t = best_j;
				if (n_mm <= 2) {
					r.mqe = r.score, r.mqe_t = qlen - 1;
					if (r.mqe + jb.end_bonus >= r.max) r.reach_end = 1, kind = 1;
				}
			} else {
				const int32_t max_gapped = (qlen - 2) * a.sc_a - 2 * (a.q + a.e);
				if (n_mm <= 3 || r.score > max_gapped) kind = 1;
			}
		}
		DpJob d{};
		d.cigar_off = jb.cigar_off;
		if (kind == 1) {
			r.n_cigar = 1;                                             // ksw_push_cigar of M qlen
			if (lane == 0) a.cigar[jb.cigar_off] = (uint32_t)qlen << 4;
		} else {
			int w = jb.w;
			if (n_mm >= 0 && (qlen < tlen ? qlen : tlen) < 100) {      // align.c:364-368
				const int32_t x = n_mm * (a.sc_a + a.sc_b);
				int32_t max1 = 0, max2 = 0;
				if (x >= a.q + a.e) max1 = (x - a.q + a.e - 1) / a.e;
				if (x >= a.q2 + a.e2) max2 = (x - a.q2 + a.e2 - 1) / a.e2;
				const int32_t mbw = max1 > max2 ? max1 : max2;
				if (w > mbw + 4) w = mbw + 4;
			}
			if (a.max_sw_mat > 0 && (int64_t)tlen * qlen > a.max_sw_mat) {   // align.c:370-372
				reset_res(r);
				r.zdropped = 1, kind = 2;
			} else {
				d.qoff = jb.qoff, d.toff = jb.toff, d.qlen = qlen, d.tlen = tlen;
				d.q = (int8_t)a.q, d.e = (int8_t)a.e, d.q2 = (int8_t)a.q2, d.e2 = (int8_t)a.e2;
				d.w = w, d.zdrop = jb.zdrop * a.sc_a, d.end_bonus = jb.end_bonus, d.flag = flag;
			}
		}
		if (lane == 0) {
			a.dp[i] = d;
			a.kind[i] = (uint8_t)kind;
			if (kind) a.res[i] = r;
		}
	}
}

// results decided before the DP replace what the DP kernels wrote for their empty jobs
__global__ void align_merge_kernel(const uint8_t *kind, const DpRes *decided, DpRes *res, int64_t n)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
		if (kind[i]) res[i] = decided[i];
}

hipError_t launch_align_plan(const AlignArgs &a, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	int64_t g = (a.n + 3) / 4;
	if (g > 8192) g = 8192;
	hipLaunchKernelGGL(align_plan_kernel, dim3((unsigned)g), dim3(256), 0, st, a);
	return hipGetLastError();
}

hipError_t launch_align_merge(const uint8_t *kind, const DpRes *decided, DpRes *res, int64_t n, hipStream_t st)
{
	if (n <= 0) return hipSuccess;
	int64_t g = (n + 255) / 256;
	if (g > 4096) g = 4096;
	hipLaunchKernelGGL(align_merge_kernel, dim3((unsigned)g), dim3(256), 0, st, kind, decided, res, n);
	return hipGetLastError();
}

} // namespace mbg
