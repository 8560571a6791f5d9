// LBVH over faces or edges: build (global box, Morton, radix sort, Karras),
// fused leaf update + bottom-up refit, and stack traversal queries.
//
// Mirrors lcs::LBVH (Solver/CollisionDetector/lbvh.h:164-230).  Node numbering
// follows the reference: inner nodes 0..N-2, leaves N-1..2N-2, so that the
// children / parent / object-index arrays can be compared with the reference's
// known answers.  Traversal itself uses a separate 64 B record per inner node
// that holds both child boxes and child references (leaf references carry the
// primitive id with the top bit set), so a visit is one contiguous load.
#pragma once

#include "ut_runtime.hpp"
#include "radix_sort.hpp"

namespace ut
{

enum TreeType : int
{
	kTreeFace = 1,
	kTreeEdge = 2,
};

// Query kinds (lbvh.cpp:780-894).
enum QueryKind : int
{
	kQueryVerts = 0,		  // fn_query_from_verts_v2: box of x_begin/x_end of a vertex
	kQueryEdges = 1,		  // fn_query_from_edges_v2: edge box, keeps eid < adj_eid
	kQueryEdgesUnculled = 2,  // fn_query_from_edges_unculled_v2: edge box, keeps all
};

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct BvhNode
{
	// floats 0..5: child 0 lo/hi, floats 6..11: child 1 lo/hi, words 12/13: child refs.
	float4 a, b, c;
	uint4  d;
};

struct Lbvh
{
	Context* ctx = nullptr;
	int		 type = kTreeFace;
	uint32_t num_leaves = 0;

	// build scratch
	DevBuf<float>	 leaf_center; // [3N]
	DevBuf<float>	 block_box;	  // [6 * blocks]
	DevBuf<float>	 global_box;  // [6]
	DevBuf<uint32_t> sort_keys[2], sort_vals[2], sort_hist;
	RadixSortBuffers sort;

	// tree (reference numbering)
	DevBuf<uint64_t> morton_sorted;		  // [N]  (morton30 << 32 | original index)
	DevBuf<uint32_t> sorted_get_original; // [N]
	DevBuf<uint32_t> parent;			  // [2N-1]
	DevBuf<uint32_t> children;			  // [2(2N-1)]
	DevBuf<uint32_t> object_idx;		  // [2N-1]
	// traversal records and refit bookkeeping
	DevBuf<BvhNode>	 nodes;			  // [max(N-1,1)]
	DevBuf<uint32_t> leaf_parent_slot;	// [N]   parent << 2 | parent-is-block-local << 1 | slot
	DevBuf<uint32_t> inner_parent_slot; // [N-1] same encoding, kNone for the root
	DevBuf<uint32_t> refit_flag;		// [N-1] arrivals per inner node, 0 between refits
	DevBuf<uint32_t> refit_exits;		// nodes reached from below whose range spans refit blocks
	DevBuf<uint32_t> refit_exit_count;
	DevBuf<float>	 root_box;			// [6]
	DevBuf<uint32_t> status;			// [2] : [0] health (1 ok), [1] truncated queries
	DevBuf<uint32_t> spill;				// traversal stack entries beyond the LDS part
	DevBuf<uint2>	 pair_buf;			// swept pairs of a CCD query (see contact_query)
	DevBuf<uint32_t> pair_count;		// [1]

	void init(Context* c, int tree_type, uint32_t n);
	// lbvh.cpp:1003-1021 (reduce_{face,edge}_tree_aabb) + fn_compute_mortons
	void reduce_aabb(const float* x, const uint32_t* prims);
	// lbvh.cpp:1023-1047: sort + fn_apply_sorted + fn_build_inner_nodes
	void construct_tree();
	// Karras build over keys already sorted in sort.keys/vals[sort.result].
	void construct_tree_presorted();
	// lbvh.cpp:621-758 + 1050-1092: leaf swept box + thickness, then refit
	void update_leaves_and_refit(const float* thickness, const float* x_start, const float* x_end, const uint32_t* prims);
	// lbvh.cpp:780-894, 1193-1262: returns nothing; count (device) may exceed max_count
	void query(int kind,
		const float*	x_begin,
		const float*	x_end,
		const uint32_t* edges,
		uint32_t		num_queries,
		const float*	d_hat,
		const float*	thickness,
		uint32_t*		count,
		uint32_t*		list,
		uint32_t		max_count);
};

} // namespace ut
