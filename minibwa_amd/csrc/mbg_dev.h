// mbg_dev.h -- device-side FM-index primitives for gfx950 (rank, extend, LF), shared by the kernels.
//
// The index stays in the reference's byte layout in HBM (bwt.c:60-65): one 64-byte block per 128
// BWT rows = 4 x u64 cumulative counts (low 56 bits: count before the block; high 8 bits: count in
// the block's first 64 symbols) + 4 x u64 packed 2-bit symbols (symbol i at bits 2*(i&31) of word
// i>>5).  A block is fetched as four 16-byte loads; counts within it use popcount on 2-bit lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbg {

struct Sai { uint64_t x0, x1, size, info; };

struct Idx {
	const uint64_t *data;   // occ blocks
	uint64_t data_len;      // in u64 words
	const uint64_t *sa;     // sampled SA
	const uint64_t *sa_fine;   // SA of every 2^fine_bit-th row, from the same walk (nullptr if absent)
	const Sai *pre;         // k-mer interval cache (nullptr if pre_len == 0)
	uint64_t primary, seq_len;
	uint64_t L2[5];
	uint32_t sa_bit, fine_bit;
	int32_t pre_len;
};

constexpr uint64_t kM56 = (1ULL << 56) - 1;

template <typename T>
__device__ __forceinline__ T pick4(int c, T a0, T a1, T a2, T a3) { return c == 0 ? a0 : c == 1 ? a1 : c == 2 ? a2 : a3; }

__device__ __forceinline__ void load_block(const uint64_t *data, uint64_t blk, uint64_t b[8])
{
	const ulonglong2 *p = reinterpret_cast<const ulonglong2*>(data + (blk << 3));
	ulonglong2 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3];
	b[0] = a0.x; b[1] = a0.y; b[2] = a1.x; b[3] = a1.y;
	b[4] = a2.x; b[5] = a2.y; b[6] = a3.x; b[7] = a3.y;
}

// number of symbol-c 2-bit lanes among the first n (0..32) lanes of w
__device__ __forceinline__ uint32_t sym_cnt(uint64_t w, uint64_t pat, int n)
{
	uint64_t e = ~(w ^ pat);
	e &= (e >> 1) & 0x5555555555555555ULL;
	uint64_t m = n >= 32 ? ~0ULL : ((1ULL << (2 * n)) - 1);
	return (uint32_t)__popcll(e & m);
}

// four per-base counts as scalars (arrays indexed at run time would go through scratch memory)
struct C4 { uint64_t a, c, g, t; };

// occurrences of each base among $-removed symbols [blk*128, blk*128 + in] plus everything
// before the block (bwt.c:160-193: the 56/8-bit counts and the half-block start)
__device__ __forceinline__ C4 occ_incl(const uint64_t b[8], uint32_t in)
{
	const int half = in >= 64;
	const int n = (int)(in & 63) + 1;
	const uint64_t w0 = half ? b[6] : b[4], w1 = half ? b[7] : b[5];
	const int n0 = n > 32 ? 32 : n, n1 = n > 32 ? n - 32 : 0;
	const uint32_t c1 = sym_cnt(w0, 0x5555555555555555ULL, n0) + sym_cnt(w1, 0x5555555555555555ULL, n1);
	const uint32_t c2 = sym_cnt(w0, 0xAAAAAAAAAAAAAAAAULL, n0) + sym_cnt(w1, 0xAAAAAAAAAAAAAAAAULL, n1);
	const uint32_t c3 = sym_cnt(w0, ~0ULL, n0) + sym_cnt(w1, ~0ULL, n1);
	const uint32_t c0 = (uint32_t)n - c1 - c2 - c3;
	C4 r;
	r.a = (b[0] & kM56) + (half ? (b[0] >> 56) : 0) + c0;
	r.c = (b[1] & kM56) + (half ? (b[1] >> 56) : 0) + c1;
	r.g = (b[2] & kM56) + (half ? (b[2] >> 56) : 0) + c2;
	r.t = (b[3] & kM56) + (half ? (b[3] >> 56) : 0) + c3;
	return r;
}

// symbol at $-removed index k (bwt.c:119); reads past the array give 0
__device__ __forceinline__ int bwt_b0(const Idx &ix, uint64_t k)
{
	uint64_t wi = (k >> 7 << 3) + 4 + ((k & 127) >> 5);
	if (wi >= ix.data_len) return 0;
	return (int)(ix.data[wi] >> ((k & 31) << 1) & 3);
}

// mb_bwt_rank2a (bwt.c:195-231) for k <= l.  One block fetch when k-1 and l-1 share a block;
// returns the number of 64-byte blocks fetched.
__device__ __forceinline__ int rank2a(const Idx &ix, uint64_t k, uint64_t l, C4 &ck, C4 &cl)
{
	int nblk = 0;
	uint64_t b[8];
	uint64_t blk = ~0ULL;
	ck = C4{0, 0, 0, 0};
	if (k != 0) {
		uint64_t j = k - 1;
		j -= (j >= ix.primary);
		blk = j >> 7;
		load_block(ix.data, blk, b);
		++nblk;
		ck = occ_incl(b, (uint32_t)(j & 127));
	}
	// bwt.c:205-209: with l==k+1 the reference adds the symbol at $-removed index k-(k>primary);
	// at k==primary that row is '$' and the shortcut's count differs from rank1a(l) -- keep it.
	if (l == k + 1 && k != 0 && k == ix.primary) {
		const int b0 = bwt_b0(ix, k);
		cl = C4{ck.a + (b0 == 0), ck.c + (b0 == 1), ck.g + (b0 == 2), ck.t + (b0 == 3)};
		return nblk;
	}
	cl = C4{0, 0, 0, 0};
	if (l != 0) {
		uint64_t j = l - 1;
		j -= (j >= ix.primary);
		if ((j >> 7) != blk) load_block(ix.data, j >> 7, b), ++nblk;
		cl = occ_incl(b, (uint32_t)(j & 127));
	}
	return nblk;
}

// array form for the probe kernels
__device__ __forceinline__ int rank2a(const Idx &ix, uint64_t k, uint64_t l, uint64_t ck[4], uint64_t cl[4])
{
	C4 a, b;
	const int n = rank2a(ix, k, l, a, b);
	ck[0] = a.a; ck[1] = a.c; ck[2] = a.g; ck[3] = a.t;
	cl[0] = b.a; cl[1] = b.c; cl[2] = b.g; cl[3] = b.t;
	return n;
}

// mb_bwt_set_intv (bwt.h:75-81)
__device__ __forceinline__ Sai set_intv(const Idx &ix, int c)
{
	Sai p;
	p.x0 = pick4(c, ix.L2[0], ix.L2[1], ix.L2[2], ix.L2[3]) + 1;
	p.x1 = pick4(c, ix.L2[3], ix.L2[2], ix.L2[1], ix.L2[0]) + 1;
	p.size = pick4(c, ix.L2[1] - ix.L2[0], ix.L2[2] - ix.L2[1], ix.L2[3] - ix.L2[2], ix.L2[4] - ix.L2[3]);
	p.info = 0;
	return p;
}

// Occurrences of base c and of all bases greater than c among $-removed symbols
// [blk*128, blk*128 + in] plus everything before the block: the two sums one child of an extension
// needs (its own rank, and the sizes of the children ordered before it on the other strand).
struct CG { uint64_t c, gt; };

__device__ __forceinline__ CG occ_cg(const uint64_t b[8], uint32_t in, int c)
{
	constexpr uint64_t K5 = 0x5555555555555555ULL;
	const int half = in >= 64;
	const int n = (int)(in & 63) + 1;
	const uint64_t w0 = half ? b[6] : b[4], w1 = half ? b[7] : b[5];
	const int n0 = n > 32 ? 32 : n, n1 = n > 32 ? n - 32 : 0;
	const uint64_t m0 = n0 >= 32 ? ~0ULL : ((1ULL << (2 * n0)) - 1), m1 = n1 >= 32 ? ~0ULL : ((1ULL << (2 * n1)) - 1);
	const uint64_t pat = (uint64_t)c * K5;
	uint64_t e0 = ~(w0 ^ pat), e1 = ~(w1 ^ pat);
	e0 &= e0 >> 1, e1 &= e1 >> 1;
	const uint64_t h0 = w0 >> 1 & K5, l0 = w0 & K5, h1 = w1 >> 1 & K5, l1 = w1 & K5;
	const uint64_t g0 = c == 0 ? (h0 | l0) : c == 1 ? h0 : c == 2 ? (h0 & l0) : 0;
	const uint64_t g1 = c == 0 ? (h1 | l1) : c == 1 ? h1 : c == 2 ? (h1 & l1) : 0;
	const uint32_t nc = (uint32_t)(__popcll(e0 & K5 & m0) + __popcll(e1 & K5 & m1));
	const uint32_t ng = (uint32_t)(__popcll(g0 & m0) + __popcll(g1 & m1));
	const uint64_t p1 = (b[1] & kM56) + (half ? (b[1] >> 56) : 0), p2 = (b[2] & kM56) + (half ? (b[2] >> 56) : 0);
	const uint64_t p3 = (b[3] & kM56) + (half ? (b[3] >> 56) : 0);
	const uint64_t p0 = (b[0] & kM56) + (half ? (b[0] >> 56) : 0);
	CG r;
	r.c = pick4(c, p0, p1, p2, p3) + nc;
	r.gt = (c == 0 ? p1 + p2 + p3 : c == 1 ? p2 + p3 : c == 2 ? p3 : 0) + ng;
	return r;
}

// rank2a restricted to (base c, bases > c); the same block fetches and the same l-k==1 quirk at
// k == primary (bwt.c:205-209) as rank2a above
__device__ __forceinline__ int rank_cg(const Idx &ix, uint64_t k, uint64_t l, int c, CG &gk, CG &gl)
{
	int nblk = 0;
	uint64_t b[8];
	uint64_t blk = ~0ULL;
	gk = CG{0, 0};
	if (k != 0) {
		uint64_t j = k - 1;
		j -= (j >= ix.primary);
		blk = j >> 7;
		load_block(ix.data, blk, b);
		++nblk;
		gk = occ_cg(b, (uint32_t)(j & 127), c);
	}
	if (l == k + 1 && k != 0 && k == ix.primary) {
		const int b0 = bwt_b0(ix, k);
		gl = CG{gk.c + (b0 == c), gk.gt + (b0 > c)};
		return nblk;
	}
	gl = CG{0, 0};
	if (l != 0) {
		uint64_t j = l - 1;
		j -= (j >= ix.primary);
		if ((j >> 7) != blk) load_block(ix.data, j >> 7, b), ++nblk;
		gl = occ_cg(b, (uint32_t)(j & 127), c);
	}
	return nblk;
}

// mb_bwt_extend (bwt.c:237-250) restricted to the one base c the caller keeps (selects instead of
// indexing, so nothing is addressed through scratch memory).  The cumulative x[is_back] of child c
// is x[is_back] + [$ in range] + the sizes of the children > c, i.e. the rank difference of "> c".
__device__ __forceinline__ Sai extend1(const Idx &ix, const Sai &ik, int is_back, int c, int *nblk = nullptr)
{
	const uint64_t xf = is_back ? ik.x0 : ik.x1;   // x[!is_back]
	const uint64_t xg = is_back ? ik.x1 : ik.x0;   // x[is_back]
	CG tk, tl;
	const int nb = rank_cg(ix, xf, xf + ik.size, c, tk, tl);
	if (nblk) *nblk += nb;
	const uint64_t acc = xg + (uint64_t)(xf <= ix.primary && xf + ik.size > ix.primary) + (tl.gt - tk.gt);
	Sai o;
	const uint64_t nf = pick4(c, ix.L2[0], ix.L2[1], ix.L2[2], ix.L2[3]) + 1 + tk.c;
	o.x0 = is_back ? nf : acc;
	o.x1 = is_back ? acc : nf;
	o.size = tl.c - tk.c;
	o.info = 0;
	return o;
}

// all four extensions (mb_bwt_extend)
__device__ __forceinline__ void extend4(const Idx &ix, const Sai &ik, int is_back, Sai ok[4])
{
	const uint64_t xf = is_back ? ik.x0 : ik.x1;
	const uint64_t xg = is_back ? ik.x1 : ik.x0;
	uint64_t tk[4], tl[4];
	rank2a(ix, xf, xf + ik.size, tk, tl);
	uint64_t acc = xg + (uint64_t)(xf <= ix.primary && xf + ik.size > ix.primary);
	for (int c = 3; c >= 0; --c) {
		const uint64_t nf = ix.L2[c] + 1 + tk[c];
		ok[c].x0 = is_back ? nf : acc;
		ok[c].x1 = is_back ? acc : nf;
		ok[c].size = tl[c] - tk[c];
		ok[c].info = 0;
		acc += ok[c].size;
	}
}

// bwt_invPsi (bwt.c:469-475): LF of row k with one block fetch.  The count of c among rows [0,k)
// (excluding '$') equals the count among $-removed symbols [0, x) with x = k - (k > primary), and
// the symbol c = B0(x) lives in the same block.
__device__ __forceinline__ uint64_t lf(const Idx &ix, uint64_t k)
{
	if (k == ix.primary) return 0;
	const uint64_t x = k - (k > ix.primary);
	uint64_t b[8];
	load_block(ix.data, x >> 7, b);
	const uint32_t in = (uint32_t)(x & 127);
	const int c = (int)(pick4((int)(in >> 5), b[4], b[5], b[6], b[7]) >> ((in & 31) << 1) & 3);
	const int half = in >= 64;
	const int n = (int)(in & 63);                 // symbols before x within the half
	const uint64_t pat = c == 0 ? 0 : c == 1 ? 0x5555555555555555ULL : c == 2 ? 0xAAAAAAAAAAAAAAAAULL : ~0ULL;
	const uint64_t w0 = half ? b[6] : b[4], w1 = half ? b[7] : b[5];
	const int n0 = n > 32 ? 32 : n, n1 = n > 32 ? n - 32 : 0;
	const uint64_t bc = pick4(c, b[0], b[1], b[2], b[3]);
	const uint64_t cnt = (bc & kM56) + (half ? (bc >> 56) : 0) + sym_cnt(w0, pat, n0) + sym_cnt(w1, pat, n1);
	return pick4(c, ix.L2[0], ix.L2[1], ix.L2[2], ix.L2[3]) + 1 + cnt;
}

// exclusive prefix sum across the 64 lanes of a converged wave; *total = the wave's sum
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t *total)
{
	const uint32_t lane = __lane_id();
	uint32_t x = v;
	#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t y = __shfl_up(x, o);
		if (lane >= (uint32_t)o) x += y;
	}
	*total = __shfl(x, 63);
	return x - v;
}

// Work distribution with one global atomic per kChunk items per wave.  The wave holds a chunk
// [next, end) of item indices (the same values in every live lane) and hands them to its lanes in
// lane order; only when the chunk runs out does its first live lane take a new chunk from the
// global counter.  Indices past the caller's item count mean "no more work".
struct WaveQueue {
	static constexpr uint32_t kChunk = 64;
	uint32_t next = 0, end = 0;
	__device__ __forceinline__ uint32_t fetch(uint32_t *counter, bool want)
	{
		const uint64_t mask = __ballot(want);
		if (mask == 0) return 0;
		const uint32_t lane = __lane_id();
		const uint32_t k = (uint32_t)__popcll(mask);
		const uint32_t r = (uint32_t)__popcll(mask & ((1ULL << lane) - 1));
		const uint32_t avail = end - next;
		uint32_t t;
		if (k <= avail) {
			t = next + r;
			next += k;
		} else {
			const uint64_t live = __ballot(1);
			const int leader = __ffsll((unsigned long long)live) - 1;
			uint32_t base = 0;
			if ((int)lane == leader) base = atomicAdd(counter, kChunk);
			base = __shfl(base, leader);
			t = r < avail ? next + r : base + (r - avail);
			next = base + (k - avail);
			end = base + kChunk;
		}
		return t;
	}
};

} // namespace mbg
