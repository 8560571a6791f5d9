// Device helpers for the gfx950 untangling kernels: XCD-aware workgroup order
// and 64-lane wavefront reductions (CDNA4), on top of the host/device math of
// ut_math.hpp.
#pragma once

#include "ut_math.hpp"

namespace ut
{

constexpr int kWave = 64;

// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs
// (workgroup i runs on XCD i % 8), and each XCD has its own L2.
constexpr uint32_t kNumXcd = 8;

// Logical workgroup id of a persistent grid (size a multiple of kNumXcd) that
// numbers the workgroups of one XCD consecutively.  A grid-stride loop over
// logical ids keeps its load balance (every round of gridDim.x tiles is shared
// by all XCDs), while the tiles one XCD works on at the same time are
// neighbours in the work order and so share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block()
{
	return (blockIdx.x % kNumXcd) * (gridDim.x / kNumXcd) + blockIdx.x / kNumXcd;
}

// Wave-wide (64 lane) reductions through cross-lane shuffles.
__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v = fminf(v, __shfl_xor(v, o, kWave));
	return v;
}
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v = fmaxf(v, __shfl_xor(v, o, kWave));
	return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o, kWave);
	return v;
}
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o, kWave);
	return v;
}
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Wave-aggregated append: every active lane with `want` gets a unique slot in
// a global list.  One atomic per wave instead of one per lane.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want)
{
	const uint64_t mask = __ballot(want);
	uint32_t	   base = 0;
	const int	   leader = __ffsll((unsigned long long)mask) - 1;
	if (mask != 0 && (int)__lane_id() == leader)
		base = atomicAdd(counter, (uint32_t)__popcll(mask));
	base = __shfl(base, leader < 0 ? 0 : leader, kWave);
	return base + (uint32_t)__popcll(mask & lanemask_lt());
}

} // namespace ut
