// Contact filtering around intersection contours: the record decode, the
// invalid-vertex test, the region match and the region entries of PRP hits and
// EF pairs, written once for the device kernels (regions.hip, the CCD broad
// phase of contact.hip) and for host callers.
//
// fn_is_matched_invalid_region (intersection_resolver.cpp:2233-2345),
// fn_identify_invalid_verts_from_proximity_pairs (:2353-2374),
// fn_filter_invalid_proximity_pairs_* (:2376-2417),
// fn_filter_broadphase_{vf,ee}_pairs (:2427-2535) and the region sets of
// host_apply_untangling_constraint (intersection_resolver2.cpp:2296-2426).
//
// Arrays are read through a template parameter A (anything with operator[]
// and a test for presence): the kernels pass raw device pointers, a host check
// can pass views that verify every index against the array's length.
#pragma once

#include "contact_rec.hpp"
#include "ut_math.hpp"

namespace ut
{

template <typename A>
struct RegionViewT
{
	A off{};	 // [V + 1], absent when no regions
	A regions{}; // region ids of each vertex (2c + side), any order
	A invalid{}; // [V] vert_is_invalid, absent when unused
};
using RegionView = RegionViewT<const uint32_t*>;

// A contact record's vertices with the type bit masked off, its two sides
// (VF: vertex | face; EE: edge | edge) and the weights of its difference
// vector x_a - x_b.
struct RecordSides
{
	uint32_t id[4];
	uint32_t a[3], b[3];
	int		 na, nb;
	bool	 vf;
	float	 w[4];
};

UT_HD RecordSides decode_record(const ContactRec& r)
{
	RecordSides s;
	s.vf = (r.idx[2] & kTypeBit) != 0u;
	for (int k = 0; k < 4; ++k)
		s.id[k] = r.idx[k] & kIndexMask;
	if (s.vf)
	{
		s.a[0] = s.id[0], s.na = 1;
		s.b[0] = s.id[1], s.b[1] = s.id[2], s.b[2] = s.id[3], s.nb = 3;
	}
	else
	{
		s.a[0] = s.id[0], s.a[1] = s.id[1], s.na = 2;
		s.b[0] = s.id[2], s.b[1] = s.id[3], s.nb = 2;
	}
	const float b0 = r.bary[0], b1 = r.bary[1];
	s.w[0] = s.vf ? 1.0f : b0;
	s.w[1] = s.vf ? -b0 : 1.0f - b0;
	s.w[2] = -b1;
	s.w[3] = s.vf ? b0 + b1 - 1.0f : b1 - 1.0f;
	return s;
}

template <typename F>
UT_HD f3 load3(const F& p, uint32_t i)
{
	return f3{ p[3 * i + 0], p[3 * i + 1], p[3 * i + 2] };
}

// fn_identify_invalid_verts_from_proximity_pairs: the record's points are
// closer than the contact offset of its first vertices (all four vertices of
// such a record become invalid).
template <typename F>
UT_HD bool record_too_close(const RecordSides& s, const F& x, const F& offset)
{
	f3 d = mk3(0.0f, 0.0f, 0.0f);
	for (int k = 0; k < 4; ++k)
		d = d + load3(x, s.id[k]) * s.w[k];
	const float t = offset[s.id[0]] + offset[s.id[2]];
	return dot3(d, d) < t * t + 1e-10f;
}

// Some region of an a-side vertex and some region of a b-side vertex are the
// two sides 2c, 2c + 1 of one contour c.
template <typename A>
UT_HD bool region_sets_match(const RegionViewT<A>& r, const uint32_t* a, int na, const uint32_t* b, int nb)
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
template <typename A>
UT_HD bool pair_filtered(const RegionViewT<A>& r, const uint32_t* a, int na, const uint32_t* b, int nb)
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

// The two sides of a swept broad-phase pair: kind 2 = (vertex q, face leaf),
// kind 3 = (edge q, edge leaf).
template <typename U>
UT_HD void ccd_pair_sides(int kind, uint32_t q, uint32_t leaf, const U& edges, const U& faces, uint32_t* s, int* ns,
	uint32_t* t, int* nt)
{
	if (kind == 2)
	{
		s[0] = q, *ns = 1;
		t[0] = faces[3 * leaf], t[1] = faces[3 * leaf + 1], t[2] = faces[3 * leaf + 2], *nt = 3;
	}
	else
	{
		s[0] = edges[2 * q], s[1] = edges[2 * q + 1], *ns = 2;
		t[0] = edges[2 * leaf], t[1] = edges[2 * leaf + 1], *nt = 2;
	}
}

// Vertices of one side of a PRP hit (type 0 VF: source vertex, target face;
// 1 EE: edges; 2 FV: source face, target vertex).  Source vertices take region
// 2c, target vertices 2c + 1.
template <typename U>
UT_HD uint32_t hit_side_verts(const U& edges, const U& faces, uint32_t type, uint32_t prim, bool src, uint32_t* o)
{
	const bool vert = (type == 0 && src) || (type == 2 && !src);
	if (vert)
		return o[0] = prim, 1u;
	if (type == 1)
		return o[0] = edges[2 * prim], o[1] = edges[2 * prim + 1], 2u;
	o[0] = faces[3 * prim], o[1] = faces[3 * prim + 1], o[2] = faces[3 * prim + 2];
	return 3u;
}

// Region of vertex k (0, 1: edge; 2..4: face) of an EF pair with region
// er = 2 * contour + mesh index.
UT_HD uint32_t ef_vertex_region(int k, uint32_t er) { return k < 2 ? er : er ^ 1u; }

// Duplicate removal in one vertex's region list v[0, n), in place: a vertex
// collects one entry per EF pair and per PRP hit that touches it, mostly
// copies of one or two regions.  The first kRegionSet distinct regions are
// kept once each (checked against a register set); entries beyond them are
// kept as they come.  The kept list has the same set of regions.  Returns the
// kept count.
constexpr int kRegionSet = 8;
UT_HD uint32_t region_dedup(uint32_t* v, uint32_t n)
{
	uint32_t set[kRegionSet];
	int		 k = 0;
	uint32_t w = 0;
	for (uint32_t i = 0; i < n; ++i)
	{
		const uint32_t r = v[i];
		bool		   seen = false;
		for (int j = 0; j < k; ++j)
			seen = seen || set[j] == r;
		if (seen)
			continue;
		if (k < kRegionSet)
			set[k++] = r;
		v[w++] = r;
	}
	return w;
}

} // namespace ut
