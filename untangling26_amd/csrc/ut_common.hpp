// Shared device/host helpers for the gfx950 untangling kernels.
//
// Vectors are plain structs of floats so that kernels read packed float[3*V]
// position arrays (12 B per vertex) straight from HBM without padding.
// All kernels are written for 64-lane wavefronts (CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define UT_HD __host__ __device__ __forceinline__

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

struct f3
{
	float x, y, z;
};

UT_HD f3 mk3(float x, float y, float z) { return f3{ x, y, z }; }
UT_HD f3 operator+(f3 a, f3 b) { return f3{ a.x + b.x, a.y + b.y, a.z + b.z }; }
UT_HD f3 operator-(f3 a, f3 b) { return f3{ a.x - b.x, a.y - b.y, a.z - b.z }; }
UT_HD f3 operator-(f3 a) { return f3{ -a.x, -a.y, -a.z }; }
UT_HD f3 operator*(f3 a, float s) { return f3{ a.x * s, a.y * s, a.z * s }; }
UT_HD f3 operator*(float s, f3 a) { return f3{ a.x * s, a.y * s, a.z * s }; }
UT_HD f3 mul3(f3 a, f3 b) { return f3{ a.x * b.x, a.y * b.y, a.z * b.z }; }
UT_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
UT_HD f3 cross3(f3 a, f3 b)
{
	return f3{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
UT_HD f3 min3(f3 a, f3 b) { return f3{ fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z) }; }
UT_HD f3 max3(f3 a, f3 b) { return f3{ fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z) }; }

UT_HD f3 ld3(const float* __restrict__ p, uint32_t i)
{
	return f3{ p[3 * i + 0], p[3 * i + 1], p[3 * i + 2] };
}
UT_HD void st3(float* __restrict__ p, uint32_t i, f3 v)
{
	p[3 * i + 0] = v.x;
	p[3 * i + 1] = v.y;
	p[3 * i + 2] = v.z;
}

struct d3
{
	double x, y, z;
};
UT_HD d3 operator+(d3 a, d3 b) { return d3{ a.x + b.x, a.y + b.y, a.z + b.z }; }
UT_HD d3 operator-(d3 a, d3 b) { return d3{ a.x - b.x, a.y - b.y, a.z - b.z }; }
UT_HD d3 operator*(d3 a, double s) { return d3{ a.x * s, a.y * s, a.z * s }; }
UT_HD d3 operator*(double s, d3 a) { return d3{ a.x * s, a.y * s, a.z * s }; }
UT_HD double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
UT_HD d3 cross3(d3 a, d3 b)
{
	return d3{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
UT_HD d3 to_d3(f3 a) { return d3{ (double)a.x, (double)a.y, (double)a.z }; }

// Axis-aligned box as two corners.
struct box3
{
	f3 lo, hi;
};

UT_HD box3 box_empty() { return box3{ f3{ 1e8f, 1e8f, 1e8f }, f3{ -1e8f, -1e8f, -1e8f } }; }
UT_HD box3 box_of(f3 a) { return box3{ a, a }; }
UT_HD box3 box_add(box3 b, f3 p) { return box3{ min3(b.lo, p), max3(b.hi, p) }; }
UT_HD box3 box_union(box3 a, box3 b) { return box3{ min3(a.lo, b.lo), max3(a.hi, b.hi) }; }
UT_HD box3 box_inflate(box3 b, float t)
{
	return box3{ f3{ b.lo.x - t, b.lo.y - t, b.lo.z - t }, f3{ b.hi.x + t, b.hi.y + t, b.hi.z + t } };
}
// Strict overlap, aabb.h:94-98 (both directions use strict comparisons).
UT_HD bool box_overlap(const box3& a, const box3& b)
{
	return (a.lo.x < b.hi.x) & (a.lo.y < b.hi.y) & (a.lo.z < b.hi.z) & (a.hi.x > b.lo.x) & (a.hi.y > b.lo.y)
		& (a.hi.z > b.lo.z);
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

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace ut
