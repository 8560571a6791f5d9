// Device view of the contour regions and invalid vertices used to filter
// proximity and CCD pairs (fn_is_matched_invalid_region,
// intersection_resolver.cpp:2233-2345; filter_broad_phase_pairs :2594-2638).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut
{

struct RegionView
{
	const uint32_t* off = nullptr;	   // [V + 1], nullptr when no regions
	const uint32_t* regions = nullptr; // sorted per vertex
	const uint32_t* invalid = nullptr; // [V] vert_is_invalid, nullptr when unused
};

__device__ __forceinline__ bool region_sets_match(const RegionView& r, const uint32_t* a, int na, const uint32_t* b,
	int nb)
{
	for (int i = 0; i < na; ++i)
		for (uint32_t p = r.off[a[i]]; p < r.off[a[i] + 1]; ++p)
		{
			const uint32_t ra = r.regions[p];
			for (int j = 0; j < nb; ++j)
				for (uint32_t q = r.off[b[j]]; q < r.off[b[j] + 1]; ++q)
					if (ra == (r.regions[q] ^ 1u))
						return true;
		}
	return false;
}

// true when the pair (a-side vertices, b-side vertices) is filtered out
__device__ __forceinline__ bool pair_filtered(const RegionView& r, const uint32_t* a, int na, const uint32_t* b, int nb)
{
	if (r.invalid)
	{
		for (int i = 0; i < na; ++i)
			if (r.invalid[a[i]])
				return true;
		for (int j = 0; j < nb; ++j)
			if (r.invalid[b[j]])
				return true;
	}
	return r.off && region_sets_match(r, a, na, b, nb);
}

} // namespace ut
