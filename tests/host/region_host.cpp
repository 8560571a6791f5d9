// Host run of the contact region filter's per-record code.
//
// The kernels of regions.hip (k_region_hits, k_region_ef, k_mark_invalid,
// k_disable_pairs) and the CCD broad-phase filter of contact.hip are thin
// loops around the functions of region_view.hpp.  This file runs the same
// functions on the CPU, one record at a time, over array views that check
// every index against the array's length before it is read or written, and
// counts the indices that would have left an array.  The tests compare the
// results with the oracle and require zero such indices.
//
// `unmasked` selects the decode the filter once had, which masked the
// collision-type bit on idx[2] only and so used idx[3] of EE records (bit 31
// set) as a vertex index; it exists only here, so the tests can show that the
// bounds check sees that mistake.
#include "region_view.hpp"

#include <stddef.h>
#include <stdint.h>
#include <vector>

namespace
{

long g_violations = 0;

template <typename T>
struct Checked
{
	const T* p = nullptr;
	size_t	 n = 0;
	explicit operator bool() const { return p != nullptr; }
	T		 operator[](size_t i) const
	{
		if (i >= n)
		{
			++g_violations;
			return T(0);
		}
		return p[i];
	}
};

template <typename T>
struct CheckedOut
{
	T*	   p = nullptr;
	size_t n = 0;
	void   set(size_t i, T v)
	{
		if (i >= n)
			++g_violations;
		else
			p[i] = v;
	}
	T get(size_t i) const
	{
		if (i >= n)
		{
			++g_violations;
			return T(0);
		}
		return p[i];
	}
};

using U32 = Checked<uint32_t>;
using F32 = Checked<float>;

ut::RecordSides decode(const ut::ContactRec& r, int unmasked)
{
	ut::RecordSides s = ut::decode_record(r);
	if (unmasked)
	{
		// the decode before the fix: only idx[2] masked, so the EE type bit
		// stays on idx[3]
		for (int k = 0; k < 4; ++k)
			s.id[k] = k == 2 ? (r.idx[k] & ut::kIndexMask) : r.idx[k];
		if (s.vf)
			s.a[0] = s.id[0], s.b[0] = s.id[1], s.b[1] = s.id[2], s.b[2] = s.id[3];
		else
			s.a[0] = s.id[0], s.a[1] = s.id[1], s.b[0] = s.id[2], s.b[1] = s.id[3];
	}
	return s;
}

} // namespace

extern "C"
{

long rh_violations() { return g_violations; }
uint32_t rh_dedup(uint32_t* v, uint32_t n) { return ut::region_dedup(v, n); }
void rh_reset() { g_violations = 0; }

// Region lists per vertex (the two passes of RegionFilter::build): kept PRP
// hits (rows of 8 u32: slot, type, src, dst, b0, b1, dist, root) of contours
// hit_contour[i], skipped when suppressed[contour]; then every EF pair (rows
// of 5 vertex ids) with region 2 * contour_index + mesh_index.  off[nv + 1]
// and vals[cap] are written, then each vertex's list is reduced in place by
// region_dedup to its first kept[v] entries; returns the number of entries
// before the reduction (may exceed cap).
long rh_region_lists(uint32_t nv, uint32_t nh, const uint32_t* hits, const uint32_t* hit_contour,
	const uint8_t* suppressed, uint32_t ncontours, uint32_t nef, const uint32_t* ef, const uint32_t* contour_index,
	const uint32_t* mesh_index, const uint32_t* edges, uint32_t ne, const uint32_t* faces, uint32_t nf, uint32_t* off,
	uint32_t* vals, size_t cap, uint32_t* kept)
{
	const U32					 E{ edges, 2 * (size_t)ne }, F{ faces, 3 * (size_t)nf };
	const U32					 HC{ hit_contour, nh };
	const Checked<uint8_t>		 SUP{ suppressed, ncontours };
	std::vector<uint32_t>		 cnt(nv + 1, 0);
	CheckedOut<uint32_t>		 C{ cnt.data(), nv + 1 };
	CheckedOut<uint32_t>		 V{ vals, cap };
	long						 total = 0;
	for (int pass = 0; pass < 2; ++pass)
	{
		for (uint32_t i = 0; i < nh; ++i)
		{
			const uint32_t c = HC[i];
			if (SUP[c])
				continue;
			const uint32_t* h = hits + 8 * (size_t)i;
			uint32_t		v[3];
			for (int side = 0; side < 2; ++side)
			{
				const uint32_t nvv = ut::hit_side_verts(E, F, h[1], side == 0 ? h[2] : h[3], side == 0, v);
				for (uint32_t k = 0; k < nvv; ++k)
				{
					const uint32_t slot = C.get(v[k]);
					C.set(v[k], slot + 1);
					if (pass)
						V.set(slot, 2 * c + (uint32_t)side);
				}
			}
		}
		for (uint32_t i = 0; i < nef; ++i)
		{
			const uint32_t er = 2 * contour_index[i] + mesh_index[i];
			for (int k = 0; k < 5; ++k)
			{
				const uint32_t v = ef[5 * (size_t)i + k];
				const uint32_t slot = C.get(v);
				C.set(v, slot + 1);
				if (pass)
					V.set(slot, ut::ef_vertex_region(k, er));
			}
		}
		if (pass == 0)
		{
			// exclusive scan, then the counters become write cursors
			uint32_t run = 0;
			for (uint32_t v = 0; v <= nv; ++v)
			{
				off[v] = run;
				run += cnt[v];
				cnt[v] = off[v];
			}
			total = run;
		}
	}
	if ((size_t)total <= cap)
		for (uint32_t v = 0; v < nv; ++v)
			kept[v] = ut::region_dedup(vals + off[v], off[v + 1] - off[v]);
	return total;
}

// filter_invalid_pairs over n records (rows of 16 u32): marks invalid vertices
// (invalid[nv]) and sets k1 = k2 = 0 on the filtered records in place.
void rh_filter_proximity(uint32_t nv, const float* x, const float* offset, uint32_t n, uint32_t* recs,
	const uint32_t* off, const uint32_t* regions, size_t nreg, uint32_t* invalid, int unmasked)
{
	const F32			 X{ x, 3 * (size_t)nv }, O{ offset, nv };
	CheckedOut<uint32_t> INV{ invalid, nv };
	for (uint32_t v = 0; v < nv; ++v)
		invalid[v] = 0;
	ut::ContactRec* R = reinterpret_cast<ut::ContactRec*>(recs);
	for (uint32_t i = 0; i < n; ++i)
	{
		const ut::RecordSides s = decode(R[i], unmasked);
		if (ut::record_too_close(s, X, O))
			for (int k = 0; k < 4; ++k)
				INV.set(s.id[k], 1u);
	}
	ut::RegionViewT<U32> rv;
	rv.off = U32{ off, off ? (size_t)nv + 1 : 0 };
	rv.regions = U32{ regions, nreg };
	rv.invalid = U32{ invalid, nv };
	for (uint32_t i = 0; i < n; ++i)
	{
		const ut::RecordSides s = decode(R[i], unmasked);
		if (ut::pair_filtered(rv, s.a, s.na, s.b, s.nb))
		{
			R[i].k1 = 0.0f;
			R[i].k2 = 0.0f;
		}
	}
}

// CCD broad-phase filter: keep[i] = 1 when swept pair i (query, leaf) of kind
// 2 (vertex, face) or 3 (edge, edge) goes on to the ACCD.
void rh_ccd_keep(int kind, uint32_t nb, const uint32_t* broad, const uint32_t* edges, uint32_t ne,
	const uint32_t* faces, uint32_t nf, uint32_t nv, const uint32_t* off, const uint32_t* regions, size_t nreg,
	const uint32_t* invalid, uint8_t* keep)
{
	const U32			 E{ edges, 2 * (size_t)ne }, F{ faces, 3 * (size_t)nf };
	ut::RegionViewT<U32> rv;
	rv.off = U32{ off, off ? (size_t)nv + 1 : 0 };
	rv.regions = U32{ regions, nreg };
	rv.invalid = U32{ invalid, invalid ? (size_t)nv : 0 };
	for (uint32_t i = 0; i < nb; ++i)
	{
		uint32_t s[3], t[3];
		int		 ns, nt;
		ut::ccd_pair_sides(kind, broad[2 * (size_t)i], broad[2 * (size_t)i + 1], E, F, s, &ns, t, &nt);
		keep[i] = ut::pair_filtered(rv, s, ns, t, nt) ? 0 : 1;
	}
}

} // extern "C"
