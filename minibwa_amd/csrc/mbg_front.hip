// mbg_front.hip -- grouping of a mini-batch's seeding records by read, between the SMEM kernels and
// mb_anchor on the device (the front pass of mbg_map_front).
//
// Seeding leaves its records (interval + read id) in task order.  mb_anchor (seed.c:268-339) works
// per read, on that read's intervals in any order: it sorts and deduplicates them first
// (mb_seed_sort_dedup, seed.c:109-138).  So the grouping only has to bring each read's records
// together: a count per read, an exclusive scan, and a scatter whose order within a read is free.
//
// The anchor slots of a read are sized from its raw records, duplicates included (an interval found
// by both SMEM passes counts twice), so they may exceed the reference's preallocation, which counts
// after deduplication (seed.c:290-293); only "at least" is required of them.
#include <hip/hip_runtime.h>
#include "mbg_internal.h"

namespace mbg {

__global__ void __launch_bounds__(256) grp_count_kernel(const SeedRec *recs, const int64_t *n_rec, int64_t cap, int32_t max_occ,
                                                        uint32_t *cnt, unsigned long long *slots)
{
	int64_t n = *n_rec;
	if (n > cap) n = cap;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
		const uint32_t rid = recs[i].rid;
		const uint64_t size = recs[i].size;
		atomicAdd(&cnt[rid], 1u);
		atomicAdd(&slots[rid], (unsigned long long)(size < (uint64_t)max_occ ? size : (uint64_t)max_occ));
	}
}

__global__ void __launch_bounds__(256) grp_scatter_kernel(const SeedRec *recs, const int64_t *n_rec, int64_t cap, const int64_t *uoff,
                                                          uint32_t *cursor, Sai *intv)
{
	int64_t n = *n_rec;
	if (n > cap) n = cap;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
		const SeedRec r = recs[i];
		const int64_t pos = uoff[r.rid] + atomicAdd(&cursor[r.rid], 1u);
		Sai s;
		s.x0 = r.x0, s.x1 = r.x1, s.size = r.size, s.info = r.info;
		intv[pos] = s;
	}
}

// each job's CIGAR (n_cigar[i] operations at cigar + cigar_off) moved to packed + poff[i]
__global__ void __launch_bounds__(256) cigar_pack_kernel(const DpJob *jobs, const DpRes *res, int32_t n, const int64_t *poff,
                                                         const uint32_t *cigar, uint32_t *packed)
{
	const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n || (jobs[i].flag & 0x01)) return;
	const int32_t nc = res[i].n_cigar;
	const uint32_t *src = cigar + jobs[i].cigar_off;
	uint32_t *dst = packed + poff[i];
	for (int32_t k = 0; k < nc; ++k) dst[k] = src[k];
}

hipError_t launch_cigar_pack(const DpJob *jobs, const DpRes *res, int32_t n, const int64_t *poff, const uint32_t *cigar, uint32_t *packed,
                             hipStream_t st)
{
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(cigar_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, jobs, res, n, poff, cigar, packed);
	return hipGetLastError();
}

hipError_t launch_grp_count(const SeedRec *recs, const int64_t *n_rec, int64_t cap, int32_t max_occ, uint32_t *cnt,
                            unsigned long long *slots, unsigned grid, hipStream_t st)
{
	hipLaunchKernelGGL(grp_count_kernel, dim3(grid), dim3(256), 0, st, recs, n_rec, cap, max_occ, cnt, slots);
	return hipGetLastError();
}

hipError_t launch_grp_scatter(const SeedRec *recs, const int64_t *n_rec, int64_t cap, const int64_t *uoff, uint32_t *cursor, Sai *intv,
                              unsigned grid, hipStream_t st)
{
	hipLaunchKernelGGL(grp_scatter_kernel, dim3(grid), dim3(256), 0, st, recs, n_rec, cap, uoff, cursor, intv);
	return hipGetLastError();
}

} // namespace mbg
