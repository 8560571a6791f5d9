// Segment-triangle intersection shared by the EF pass and the PRP rays:
// LineIntersection (intersection_resolver_helper.cpp:333-481).  FP32, no
// contraction (the library is built with -ffp-contract=off).
#pragma once

#include "ut_common.hpp"

namespace ut
{

// RayTriangleIntersectionAndBary with the UE default tolerances (1e-6).
__device__ __forceinline__ bool line_intersection(f3 s, f3 e, f3 A, f3 B, f3 C, f3* bary, float* t_out)
{
	const float kSmall = 1e-6f;
	const f3	se = e - s;
	const float len2 = dot3(se, se);
	if (len2 < kSmall * kSmall)
		return false;
	const float len = sqrtf(len2);
	const float inv_len = 1.0f / len;
	const f3	ray = se * inv_len;
	const f3	AB = B - A;
	const f3	AC = C - A;
	const f3	N = cross3(AB, AC);
	const f3	neg = -ray;
	const float nlen = sqrtf(dot3(N, N));
	const float den = dot3(neg, N);
	if (fabsf(den) < nlen * kSmall)
		return false;
	const float inv_den = 1.0f / den;
	const f3	ra = s - A;
	const float tm = dot3(ra, N) * inv_den;
	if (tm < 0.0f || tm > len)
		return false;
	const f3	rx = cross3(neg, ra);
	const float uu = dot3(AC, rx) * inv_den;
	if (uu < kSmall || uu > (1.0f - kSmall))
		return false;
	const float vv = -dot3(AB, rx) * inv_den;
	if (vv < kSmall || (vv + uu) > (1.0f - kSmall))
		return false;
	*bary = mk3(1.0f - uu - vv, uu, vv);
	*t_out = tm * inv_len;
	return true;
}

} // namespace ut
