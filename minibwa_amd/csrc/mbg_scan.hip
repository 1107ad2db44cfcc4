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

} // namespace mbg
