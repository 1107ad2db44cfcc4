// mbg_extra.hip -- the hit finishing of align.c on gfx950: mb_fix_cigar (align.c:70-142) and the
// scores mb_update_extra sets (align.c:219-258), one thread per hit, so that a batch's CIGARs come
// back final and its sequences need not cross PCIe for them.  The =/X, cs and MD strings are
// output formatting and stay on the host.
#include "mbg_internal.h"

namespace mbg {


__global__ void __launch_bounds__(128) extra_kernel(ExtraArgs a)
{
	// products and sums rounded one at a time, as the host code computes them (no fused multiply-add)
#pragma clang fp contract(off)
	__shared__ int8_t mat[32];
	if (threadIdx.x < 32) mat[threadIdx.x] = a.mat[threadIdx.x < 28 ? threadIdx.x : 27];
	__syncthreads();
	const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= a.n) return;
	ExtraJob h = a.jobs[j];
	uint32_t *cg = a.cigar + h.cigar_off;
	const uint8_t *qseq = a.qseq + h.qoff, *tseq = a.tseq + h.toff;
	uint32_t n = (uint32_t)h.n_cigar, k;
	if (n == 0) return;                               // no hit extra: nothing is set (align.c:225)
	int32_t toff = 0, qoff = 0, qshift = 0, tshift = 0;
	bool to_shrink = false;
	if (n > 1) {
		// indel left alignment between two M runs
		for (k = 0; k < n; ++k) {
			const uint32_t op = cg[k] & 0xf, len = cg[k] >> 4;
			if (len == 0) to_shrink = true;
			if (op == 0) {
				toff += len, qoff += len;
			} else if (op == 1 || op == 2) {
				if (k > 0 && k < n - 1 && (cg[k - 1] & 0xf) == 0 && (cg[k + 1] & 0xf) == 0) {
					const int prev = (int)(cg[k - 1] >> 4);
					const uint8_t *sq = op == 1 ? qseq : tseq;
					const int off = op == 1 ? qoff : toff;
					int l = 0;
					while (l < prev && sq[off - 1 - l] == sq[off + (int)len - 1 - l]) ++l;
					if (l > 0) cg[k - 1] -= (uint32_t)l << 4, cg[k + 1] += (uint32_t)l << 4, qoff -= l, toff -= l;
					if (l == prev) to_shrink = true;
				}
				if (op == 1) qoff += len;
				else toff += len;
			} else if (op == 3) {
				toff += len;
			}
		}
		// a run of adjacent I and D operations becomes one I and one D
		for (k = 0; k + 2 < n; ++k) {
			if ((cg[k] & 0xf) > 0 && (cg[k] & 0xf) + (cg[k + 1] & 0xf) == 3) {
				uint32_t l, ci = 0, cd = 0;
				for (l = k; l < n; ++l) {
					const uint32_t op = cg[l] & 0xf;
					if (op == 1 || op == 2 || cg[l] >> 4 == 0) {
						if (op == 1) ci += cg[l] >> 4;
						else if (op == 2) cd += cg[l] >> 4;
					} else break;
				}
				if (ci > 0 && cd > 0 && l - k > 2) {
					cg[k] = ci << 4 | 1;
					cg[k + 1] = cd << 4 | 2;
					for (k += 2; k < l; ++k) cg[k] &= 0xf;
					to_shrink = true;
				}
				k = l;
			}
		}
		if (to_shrink) {                              // drop zero-length operations, merge equal neighbours
			uint32_t l = 0;
			for (k = 0; k < n; ++k)
				if (cg[k] >> 4 != 0) cg[l++] = cg[k];
			n = l;
			for (k = l = 0; k < n; ++k)
				if (k == n - 1 || (cg[k] & 0xf) != (cg[k + 1] & 0xf)) cg[l++] = cg[k];
				else cg[k + 1] += cg[k] >> 4 << 4;
			n = l;
		}
		if ((cg[0] & 0xf) == 1 || (cg[0] & 0xf) == 2) {   // a leading I or D
			const int32_t l = (int32_t)(cg[0] >> 4);
			if ((cg[0] & 0xf) == 1) {
				if (h.rev) h.qe -= l;
				else h.qs += l;
				qshift = l;
			} else {
				h.ts += l, tshift = l;
			}
			--n;
			for (k = 0; k < n; ++k) cg[k] = cg[k + 1];
		}
	}
	qseq += qshift, tseq += tshift;
	// blen, mlen, ambiguous bases and the best prefix score (align.c:229-258)
	int32_t blen = 0, mlen = 0, n_ambi = 0;
	double sc = 0.0, best = 0.0;
	toff = qoff = 0;
	for (k = 0; k < n; ++k) {
		const uint32_t op = cg[k] & 0xf, len = cg[k] >> 4;
		int na = 0, nd = 0;
		if (op == 0) {
			for (uint32_t l = 0; l < len; ++l) {
				const int cq = qseq[qoff + l], ct = tseq[toff + l];
				const int s = mat[(ct < 5 ? ct : 4) * 5 + (cq < 5 ? cq : 4)];
				if (ct > 3 || cq > 3) ++na;
				else if (ct != cq) nd += s < 0;
				sc += s;
				if (sc < 0) sc = 0;
				else best = best > sc ? best : sc;
			}
			blen += (int32_t)len - na, mlen += (int32_t)len - (na + nd), n_ambi += na;
			toff += len, qoff += len;
		} else if (op == 1 || op == 2) {
			const uint8_t *sq = op == 1 ? qseq + qoff : tseq + toff;
			for (uint32_t l = 0; l < len; ++l) na += sq[l] > 3;
			blen += (int32_t)len - na, n_ambi += na;
			if (a.log_gap) sc -= a.q + (double)a.e * (double)mb_log2_dev((float)(1.0 + len));
			else sc -= a.q + a.e;
			if (sc < 0) sc = 0;
			if (op == 1) qoff += len;
			else toff += len;
		}
	}
	h.n_cigar = (int32_t)n, h.blen = blen, h.mlen = mlen, h.n_ambi = n_ambi, h.dp_max = (int32_t)(best + .499);
	a.jobs[j] = h;
}

hipError_t launch_extra(const ExtraArgs &a, hipStream_t st)
{
	if (a.n <= 0) return hipSuccess;
	hipLaunchKernelGGL(extra_kernel, dim3((unsigned)((a.n + 127) / 128)), dim3(128), 0, st, a);
	return hipGetLastError();
}

} // namespace mbg
