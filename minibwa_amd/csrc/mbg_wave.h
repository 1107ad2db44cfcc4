// mbg_wave.h -- 64-lane wavefront primitives shared by the DP kernels
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbg {

// Cross-lane steps through DPP (VALU data-parallel primitives) rather than LDS permutes: the row
// recurrences are latency-bound, and a DPP move costs a few cycles where ds_bpermute costs an LDS
// round trip.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
	return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

// value of lane-1 (lane 0 receives 0)
__device__ __forceinline__ int wave_shr1(int v) { return (int)dpp_u32<0x138>((uint32_t)v); }

// value of lane-1; lane 0 receives `fill`
__device__ __forceinline__ int wave_shr1_fill(int v, int fill)
{
	return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false);
}

// inclusive unsigned prefix max over the 64 lanes in lane order (0 is the identity): rows of 16
// first, then the row broadcasts fold rows 0-1 and 2-3, then lanes 31 -> 63
__device__ __forceinline__ uint32_t wave_scan_max_u32(uint32_t v)
{
	uint32_t m;
	m = dpp_u32<0x111>(v); v = m > v ? m : v;   // row_shr:1
	m = dpp_u32<0x112>(v); v = m > v ? m : v;   // row_shr:2
	m = dpp_u32<0x114>(v); v = m > v ? m : v;   // row_shr:4
	m = dpp_u32<0x118>(v); v = m > v ? m : v;   // row_shr:8
	m = dpp_u32<0x142, 0xa>(v); v = m > v ? m : v;   // row_bcast:15
	m = dpp_u32<0x143, 0xc>(v); v = m > v ? m : v;   // row_bcast:31
	return v;
}

// unsigned max over the 64 lanes, returned as a wave-uniform value
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
	return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_max_u32(v), 63);
}

} // namespace mbg
