// Intersection contours from EF pairs on the GPU.
//
// Replaces the host side of IntersectionResolver::host_preprocess_ef_pairs
// (intersection_resolver.cpp:683-1096: canonical sort, EF adjacency,
// loop-vertex pairs) and intersection_contour_construction_from_ext_adjacent /
// contour_flood_filling_template (intersection_resolver2.cpp:254-309,
// intersection_resolver_helper.cpp:518-575: components with mesh parity).
#pragma once

#include "ut_runtime.hpp"
#include "ef_dcd.hpp"

namespace ut
{

struct ContourBuilder
{
	Context*		ctx = nullptr;
	uint32_t		nv = 0, ne = 0, nf = 0;
	const uint32_t* edges = nullptr;		  // [2E]  caller's device arrays
	const uint32_t* faces = nullptr;		  // [3F]
	const uint32_t* face_adj_edges = nullptr; // [3F]  edge between face[j], face[j+1]
	const uint32_t* edge_adj_faces = nullptr; // [2E]  ascending, kNone when absent

	uint32_t num_input = 0;	   // EF pairs handed to build()
	uint32_t num_pairs = 0;	   // after the degenerate filter (before loop pairs)
	uint32_t num_total = 0;	   // + loop pairs
	uint32_t num_adj = 0;	   // adjacency entries
	uint32_t num_contours = 0;
	uint32_t cap = 0;

	DevBuf<uint32_t>  sort_k[2], sort_v[2], sort_h, scan_tmp, scalars;
	DevBuf<uint64_t>  keys;		  // sorted (eid << 32 | fid) of the non-loop pairs
	DevBuf<uint32_t>  ids;		  // [2 * total] (eid, fid), sorted, loop pairs appended
	DevBuf<EfPairRec> pairs;	  // [total]
	DevBuf<uint32_t>  loop_cnt;	  // [P]
	DevBuf<uint32_t>  loop_off;	  // [P]
	DevBuf<uint32_t>  adj;		  // [num_adj] neighbours, packed by pair
	DevBuf<uint32_t>  adj_off;	  // [total + 1] start of each pair's neighbours
	DevBuf<uint32_t>  adj_n;	  // [total] neighbour counts
	DevBuf<uint32_t>  uf;		  // [2 * total] union-find parents of the parity-doubled graph
	DevBuf<uint32_t>  comp_min;	  // [total] smallest pair index of the component
	DevBuf<uint32_t>  mesh_index; // [total] parity relative to that pair
	DevBuf<uint32_t>  is_root;	  // [total]
	DevBuf<uint32_t>  rank;		  // [total]
	DevBuf<uint32_t>  contour_index;	// [total]
	DevBuf<uint32_t>  comp_inconsistent; // [total] indexed by component min pair
	// classify_contour_type / mesh_is_closed (intersection_resolver2.cpp:311-375, :2811-2845)
	uint32_t		  cap_contours = 0;
	DevBuf<uint32_t>  contour_stats; // [8 * contours] per-contour counts
	DevBuf<uint32_t>  contour_type;	 // [contours] ContourType
	DevBuf<uint32_t>  mesh_closed;	 // [2 * contours] role 0, role 1 closed

	void init(Context* c, uint32_t nv, uint32_t ne, uint32_t nf, const uint32_t* edges, const uint32_t* faces,
		const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces);
	// num_pairs EF pairs in any order; filter_degenerate is the reference's
	// !ignore_near_zero_dist_pairs.  Synchronises to learn the counts.
	void build(uint32_t num_pairs, const EfPairRec* in_pairs, const uint32_t* in_ids, bool filter_degenerate);
	void reserve(uint32_t p);
};

// Connected components with parity of an undirected graph in CSR form (the
// interface of UnitTest/test_parallel_contour_construction.cpp:281-448):
// comp_min[v] = smallest vertex of v's component, parity[v] = parity of v
// relative to it, inconsistent[comp_min] = 1 when the component has an odd cycle.
void graph_components(Context* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
	const uint32_t* flip_w, uint32_t* uf, uint32_t* comp_min, uint32_t* parity, uint32_t* inconsistent,
	uint32_t* status);

} // namespace ut
