// Contact filtering around intersection contours.
//
// IntersectionResolver::filter_invalid_pairs (intersection_resolver.cpp:2640-2682)
// and filter_broad_phase_pairs (:2594-2638): a vertex is invalid when it is
// part of a proximity pair closer than its contact thickness; proximity pairs
// touching an invalid vertex are disabled, and so are pairs whose two sides lie
// on opposite sides of one contour (region 2c + r against 2c + (r ^ 1),
// fn_is_matched_invalid_region :2233-2345).  Regions come from the PRP response
// pairs and the EF pairs (host_apply_untangling_constraint,
// intersection_resolver2.cpp:2296-2426).  CCD skips the same pairs.
#pragma once

#include "contact.hpp"
#include "contour.hpp"
#include "prp.hpp"
#include "ut_runtime.hpp"
#include "region_view.hpp"

namespace ut
{

struct RegionFilter
{
	Context*		 ctx = nullptr;
	uint32_t		 nv = 0;
	DevBuf<uint32_t> off, regions, cnt, invalid, scan_tmp;
	DevBuf<uint32_t> off2, regions2; // the lists without duplicates (what the filters read)
	bool			 active = false;

	void init(Context* c, uint32_t nv);
	// Region CSR from the kept PRP hits (source vertices 2c, target vertices
	// 2c + 1, suppressed contours skipped) and from every EF pair.
	void build(const PrpResolver& prp, const ContourBuilder& ct);
	// vert_is_invalid from proximity pairs [0, n), then disable (k1 = k2 = 0)
	// those pairs that touch an invalid vertex or match regions.
	void filter_proximity(ContactRec* pairs, uint32_t n, const float* x, const float* offset);
	RegionView view() const;
	void	   clear() { active = false; }
};

} // namespace ut
