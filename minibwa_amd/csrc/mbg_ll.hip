// mbg_ll.hip -- ksw_ll for gfx950: local Smith-Waterman with affine gaps that reports the best
// score, its end on both sequences and the second best away from it, as the mate-rescue step
// calls it (pe.c:260-272).
//
// The reference (ksw2_ll_sse.c:49-333) is Farrar's striped SSE algorithm with 16 u8 or 8 i16
// vector lanes; its results depend on that layout, and this kernel keeps them:
//   * the query is padded to p*slen positions (p = 16 for the u8 core, 8 for the i16 core,
//     slen = ceil(qlen/p)); padded positions score -1 and take part in the row maxima;
//   * the first pass carries F only inside a stripe of slen consecutive query positions (one SSE
//     lane), and E of the next row and the row maximum come from that first-pass H;
//   * the lazy-F loop then raises H to the fixed point of the full-column F (for gapo >= 1 its
//     early exit fires only once no lane can change); that H feeds the next row's diagonal and qe;
//   * u8 values saturate; they equal exact values until a row maximum reaches 255 - shift, the
//     row at which the u8 core stops and reports 255.
//
// One 64-lane wavefront per job.  Lane l holds query positions l*K .. l*K+K-1 in registers; one
// target base is one row step.  Both F chains are max-plus scans of the sources
// Z(k) = max(D(k), E(k)) - (gapo+gape) decaying by gape per position, so with
// Y(k) = max(D(k), E(k)) + k*gape:
//     F(k) = max(0, max_{k' < k} Y(k') - gapo - k*gape)
// taken over all k' < k (full column) or over k' in k's stripe only.  The stripe-local prefix max
// keys Y with the stripe index above bit 20: a prefix max then never reaches across a stripe
// start, since keys from earlier stripes come out negative after the stripe's offset is removed.
// The two scans are lane-local prefix maxima plus one DPP wave scan each.
#include "mbg_internal.h"
#include "mbg_wave.h"

namespace mbg {

// jobs (one wave each) per block: the LDS profile of a block stays within 64 KB
template <int K> constexpr int ll_waves() { return K >= 12 ? 2 : 4; }
constexpr int32_t kLlOut = 1 << 28;      // score and F offset of slots past the padded query
constexpr int kSegShift = 20;            // stripe index position in the stripe-local scan key

__device__ __forceinline__ int imax3(int a, int b, int c) { return max(max(a, b), c); }

// LDS per wave: the query profile, m rows of K*64 int32 scores laid out [t][k][lane]; each lane
// reads back only the entries it wrote
template <int K>
__global__ __launch_bounds__(64 * ll_waves<K>()) void ll_kernel(LlArgs a)
{
	extern __shared__ int32_t ll_lds[];
	const int lane = threadIdx.x & 63;
	const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int idx = blockIdx.x * ll_waves<K>() + wv;
	if (idx >= a.n) return;
	const int jid = a.order ? a.order[idx] : idx;
	const LlJob jb = a.jobs[jid];
	const int qlen = jb.qlen, tlen = jb.tlen, m = a.m;
	const bool u8 = jb.size <= 1;
	const int p = u8 ? 16 : 8;
	const int slen = (qlen + p - 1) / p, N = slen * p;
	const int gape = a.gape, gapoe = a.gapo + a.gape;
	int32_t *prof = ll_lds + wv * (m * K * 64);
	if (a.check && (qlen < 0 || tlen < 0 || N > 64 * K || a.boff[jid] + tlen > a.rows_cap ||
	                (!u8 && (int64_t)a.qmax * min(qlen, tlen) > 32767))) {
		if (lane == 0) a.res[jid] = LlRes{-1, -1, -1, -1, -1};
		return;
	}

	// query profile (ksw2_ll_sse.c:77-94) and per-slot constants
	int cf[K], sg[K], df[K];
	const uint8_t *q = a.qseq + jb.qoff;
#pragma unroll
	for (int k = 0; k < K; ++k) {
		const int pos = lane * K + k;
		const bool valid = pos < N;
		const int c = pos < qlen ? min((int)q[pos], m - 1) : 7;
		for (int t = 0; t < m; ++t)
			prof[(t * K + k) * 64 + lane] = valid ? (int)a.sc[t * 8 + c] : -kLlOut;
		cf[k] = pos * gape;
		sg[k] = valid ? (pos / slen) << kSegShift : 0;
		df[k] = valid ? a.gapo + pos * gape : kLlOut;
	}

	int H[K], E[K];
#pragma unroll
	for (int k = 0; k < K; ++k) H[k] = 0, E[k] = 0;
	int gmax = 0, te = -1, qe = N > 0 ? 0 : -1;
	int nb = 0, brow = 0, bval = 0, nemit = 0;    // the b array (ksw2_ll_sse.c:188-193)
	uint64_t *bb = a.bbuf + a.boff[jid];
	const uint8_t *tg = a.tseq + jb.toff;
	int tv = 0, rows = tlen;
	for (int i = 0; i < rows; ++i) {
		if ((i & 63) == 0) tv = i + lane < tlen ? (int)tg[i + lane] : 0;
		const int t = min(__builtin_amdgcn_readlane(tv, i & 63), m - 1);
		const int32_t *pr = prof + t * (K * 64) + lane;
		int D[K], pf[K], ps[K];
		const int hin = wave_shr1(H[K - 1]);     // H(i-1, l*K-1); lane 0 gets H(i-1,-1) = 0
#pragma unroll
		for (int k = 0; k < K; ++k) {
			D[k] = (k == 0 ? hin : H[k - 1]) + pr[k * 64];
			const int yf = max(D[k], E[k]) + cf[k];
			const int ys = yf + sg[k];
			pf[k] = k == 0 ? yf : max(pf[k - 1], yf);
			ps[k] = k == 0 ? ys : max(ps[k - 1], ys);
		}
		const int cfin = wave_shr1((int)wave_scan_max_u32((uint32_t)pf[K - 1]));
		const int csin = wave_shr1((int)wave_scan_max_u32((uint32_t)ps[K - 1]));
		int vmax = 0;
#pragma unroll
		for (int k = 0; k < K; ++k) {
			const int ef = k == 0 ? cfin : max(cfin, pf[k - 1]);
			const int es = k == 0 ? csin : max(csin, ps[k - 1]);
			const int hp = imax3(D[k], E[k], es - sg[k] - df[k]);      // first-pass H
			H[k] = max(hp, ef - df[k]);                                 // after lazy-F
			E[k] = imax3(E[k] - gape, hp - gapoe, 0);
			vmax = max(vmax, hp);
		}
		const int imax = (int)wave_max_u32((uint32_t)vmax);
		if (imax >= a.minsc) {
			if (nb == 0 || brow + 1 != i) {
				if (nb > 0) {
					if (lane == 0) bb[nemit] = (uint64_t)bval << 32 | (uint32_t)brow;
					++nemit;
				}
				++nb, brow = i, bval = imax;
			} else if (bval < imax) {
				bval = imax, brow = i;
			}
		}
		if (imax > gmax) {
			gmax = imax, te = i;
			uint32_t cand = 0;                     // smallest position holding imax, as ~pos
#pragma unroll
			for (int k = K - 1; k >= 0; --k)
				if (H[k] == imax) cand = ~(uint32_t)(lane * K + k);
			qe = (int)~wave_max_u32(cand);
			if ((u8 && gmax + a.shift >= 255) || gmax >= a.endsc) rows = i + 1;
		}
	}

	LlRes r;
	const bool s255 = u8 && gmax + a.shift >= 255;
	r.score = s255 ? 255 : gmax, r.te = te, r.qe = -1, r.score2 = -1, r.te2 = -1;
	if (!s255) {
		r.qe = qe;
		if (nb > 0) {                              // second best (ksw2_ll_sse.c:205-213)
			if (lane == 0) bb[nemit] = (uint64_t)bval << 32 | (uint32_t)brow;
			++nemit;
			// lane 0's stores complete and this CU's L1 drops stale lines before the lanes read back
			__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
			const int L = (r.score + a.qmax - 1) / a.qmax, lo = te - L, hi = te + L;
			int bv = -1, br = -1;
			for (int e = lane; e < nemit; e += 64) {
				const uint64_t x = bb[e];
				const int row = (int)(uint32_t)x, val = (int)(x >> 32);
				if ((row < lo || row > hi) && val > bv) bv = val, br = row;
			}
			const int v2 = (int)wave_max_u32((uint32_t)(bv + 1)) - 1;
			if (v2 >= 0) {
				r.score2 = v2;
				r.te2 = (int)~wave_max_u32(bv == v2 ? ~(uint32_t)br : 0u);
			}
		}
	}
	if (lane == 0) a.res[jid] = r;
}

int ll_slots(int qlen, int size)
{
	const int p = size <= 1 ? 16 : 8, n = (qlen + p - 1) / p * p;
	if (qlen < 0 || n > kLlMaxQlen) return -1;
	static const int ks[] = {1, 2, 3, 4, 6, 8, 12, 16};
	for (int k : ks)
		if (n <= 64 * k) return k;
	return -1;
}

template <int K>
static hipError_t launch_ll_k(const LlArgs &a, hipStream_t st)
{
	constexpr int W = ll_waves<K>();
	const unsigned grid = (unsigned)((a.n + W - 1) / W);
	const size_t shm = (size_t)W * a.m * K * 64 * sizeof(int32_t);
	hipLaunchKernelGGL(ll_kernel<K>, dim3(grid), dim3(64 * W), shm, st, a);
	return hipGetLastError();
}

hipError_t launch_ll(const LlArgs &a, int slots, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	switch (slots) {
	case 1: return launch_ll_k<1>(a, st);
	case 2: return launch_ll_k<2>(a, st);
	case 3: return launch_ll_k<3>(a, st);
	case 4: return launch_ll_k<4>(a, st);
	case 6: return launch_ll_k<6>(a, st);
	case 8: return launch_ll_k<8>(a, st);
	case 12: return launch_ll_k<12>(a, st);
	case 16: return launch_ll_k<16>(a, st);
	default: return hipErrorInvalidValue;
	}
}

} // namespace mbg
