// Host/device vector and box arithmetic shared by the gfx950 kernels and the
// host-side checks.
//
// Vectors are plain structs of floats so that kernels read packed float[3*V]
// position arrays (12 B per vertex) straight from HBM without padding.  Only
// __host__ __device__ code lives here: a plain C++ compiler can include this
// header (with the HIP headers on its include path) to run the same arithmetic
// on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define UT_HD __host__ __device__ __forceinline__

namespace ut
{

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

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace ut
