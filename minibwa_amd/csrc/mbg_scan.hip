// mbg_scan.hip -- device-wide exclusive prefix sum (rocPRIM through hipCUB) used to place
// per-task outputs without atomics.
#include <hipcub/hipcub.hpp>
#include "mbg_internal.h"

namespace mbg {

// with temp == nullptr only *temp_bytes is set
hipError_t scan_u32(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, int64_t n, hipStream_t st)
{
	return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, in, out, (int)n, st);
}

namespace {
struct ToI64 {
	__host__ __device__ int64_t operator()(uint32_t x) const { return (int64_t)x; }
	__host__ __device__ int64_t operator()(unsigned long long x) const { return (int64_t)x; }
};
} // namespace

hipError_t scan_u32_i64(void *temp, size_t *temp_bytes, const uint32_t *in, int64_t *out, int64_t n, hipStream_t st)
{
	hipcub::TransformInputIterator<int64_t, ToI64, const uint32_t*> it(in, ToI64());
	return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, it, out, (int)n, st);
}

hipError_t scan_u64_i64(void *temp, size_t *temp_bytes, const unsigned long long *in, int64_t *out, int64_t n, hipStream_t st)
{
	hipcub::TransformInputIterator<int64_t, ToI64, const unsigned long long*> it(in, ToI64());
	return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, it, out, (int)n, st);
}

namespace {
struct LlRows {
	__host__ __device__ int64_t operator()(const LlJob &j) const { return j.tlen > 0 ? j.tlen : 0; }
};
} // namespace

// exclusive prefix sum of the jobs' target lengths: each job's span of the per-row maxima list
hipError_t scan_ll_rows(void *temp, size_t *temp_bytes, const LlJob *jobs, int64_t *out, int64_t n, hipStream_t st)
{
	hipcub::TransformInputIterator<int64_t, LlRows, const LlJob*> in(jobs, LlRows());
	return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, in, out, (int)n, st);
}

} // namespace mbg
