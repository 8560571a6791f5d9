// LBVH kernels for gfx950.  See lbvh.hpp for the data layout.
#include "ut_common.hpp"
#include "lbvh.hpp"
#include "lbvh_traverse.hpp"

namespace ut
{

constexpr int kBoxThreads = 256;
// Leaves handled by one refit workgroup; inner nodes whose leaf range lies in
// one such block are refit with workgroup-scope synchronisation only.
constexpr int kRefitBlock = 256;

// ---------------------------------------------------------------------------
// Morton code, lbvh.cpp:84-135.  10 bits per axis, x in the highest lane.
__device__ __forceinline__ uint32_t expand_bits10(uint32_t v)
{
	v = (v | (v << 16)) & 0x030000FFu;
	v = (v | (v << 8)) & 0x0300F00Fu;
	v = (v | (v << 4)) & 0x030C30C3u;
	return (v | (v << 2)) & 0x09249249u;
}
__device__ __forceinline__ uint32_t morton30(f3 p)
{
	const float x = fmaxf(fminf(p.x * 1024.0f, 1023.0f), 0.0f);
	const float y = fmaxf(fminf(p.y * 1024.0f, 1023.0f), 0.0f);
	const float z = fmaxf(fminf(p.z * 1024.0f, 1023.0f), 0.0f);
	return (expand_bits10((uint32_t)x) << 2) | (expand_bits10((uint32_t)y) << 1) | expand_bits10((uint32_t)z);
}

__device__ __forceinline__ void prim_points(
	int type, const float* __restrict__ x, const uint32_t* __restrict__ prims, uint32_t i, f3* p, uint32_t* v0)
{
	if (type == kTreeFace)
	{
		const uint32_t a = prims[3 * i], b = prims[3 * i + 1], c = prims[3 * i + 2];
		p[0] = ld3(x, a);
		p[1] = ld3(x, b);
		p[2] = ld3(x, c);
		*v0 = a;
	}
	else
	{
		const uint32_t a = prims[2 * i], b = prims[2 * i + 1];
		p[0] = ld3(x, a);
		p[1] = ld3(x, b);
		p[2] = p[1];
		*v0 = a;
	}
}

// Leaf centroid + per-block box of the primitive vertices (first pass of
// fn_reduce_{face,edge}_tree_global_aabb, lbvh.cpp:400-496).
__global__ __launch_bounds__(kBoxThreads) void k_leaf_center_box(int type,
	const float* __restrict__ x,
	const uint32_t* __restrict__ prims,
	uint32_t n,
	float* __restrict__ center,
	float* __restrict__ block_box)
{
	__shared__ float red[6][kBoxThreads / kWave];
	const uint32_t	 i = blockIdx.x * kBoxThreads + threadIdx.x;
	box3			 bx = box_empty();
	if (i < n)
	{
		f3		 p[3];
		uint32_t v0;
		prim_points(type, x, prims, i, p, &v0);
		f3 c;
		if (type == kTreeFace)
		{
			c = 0.333333f * ((p[0] + p[1]) + p[2]); // lbvh.cpp:479
			bx = box_add(box_add(box_of(p[0]), p[1]), p[2]);
		}
		else
		{
			c = 0.5f * (p[0] + p[1]); // lbvh.cpp:461
			bx = box_add(box_of(p[0]), p[1]);
		}
		st3(center, i, c);
	}
	float v[6] = { bx.lo.x, bx.lo.y, bx.lo.z, bx.hi.x, bx.hi.y, bx.hi.z };
#pragma unroll
	for (int k = 0; k < 3; ++k)
		v[k] = wave_min(v[k]);
#pragma unroll
	for (int k = 3; k < 6; ++k)
		v[k] = wave_max(v[k]);
	const uint32_t w = threadIdx.x / kWave;
	if (__lane_id() == 0)
	{
#pragma unroll
		for (int k = 0; k < 6; ++k)
			red[k][w] = v[k];
	}
	__syncthreads();
	if (threadIdx.x < 6)
	{
		const int k = threadIdx.x;
		float	  r = red[k][0];
		for (int q = 1; q < kBoxThreads / kWave; ++q)
			r = k < 3 ? fminf(r, red[k][q]) : fmaxf(r, red[k][q]);
		block_box[6 * blockIdx.x + k] = r;
	}
}

// Second pass of the global box (every block folds all block boxes; they sit
// in L2) followed by fn_compute_mortons, lbvh.cpp:512-528.
__global__ __launch_bounds__(256) void k_morton(uint32_t n,
	const float* __restrict__ center,
	const float* __restrict__ block_box,
	uint32_t num_blocks,
	float* __restrict__ global_box,
	uint32_t* __restrict__ keys,
	uint32_t* __restrict__ vals)
{
	__shared__ float red[6][256 / kWave];
	__shared__ float gb[6];
	float			 r[6] = { 1e8f, 1e8f, 1e8f, -1e8f, -1e8f, -1e8f };
	for (uint32_t b = threadIdx.x; b < num_blocks; b += 256)
	{
#pragma unroll
		for (int k = 0; k < 3; ++k)
			r[k] = fminf(r[k], block_box[6 * b + k]);
#pragma unroll
		for (int k = 3; k < 6; ++k)
			r[k] = fmaxf(r[k], block_box[6 * b + k]);
	}
#pragma unroll
	for (int k = 0; k < 3; ++k)
		r[k] = wave_min(r[k]);
#pragma unroll
	for (int k = 3; k < 6; ++k)
		r[k] = wave_max(r[k]);
	const uint32_t w = threadIdx.x / kWave;
	if (__lane_id() == 0)
	{
#pragma unroll
		for (int k = 0; k < 6; ++k)
			red[k][w] = r[k];
	}
	__syncthreads();
	if (threadIdx.x < 6)
	{
		const int k = threadIdx.x;
		float	  v = red[k][0];
		for (int q = 1; q < 256 / kWave; ++q)
			v = k < 3 ? fminf(v, red[k][q]) : fmaxf(v, red[k][q]);
		gb[k] = v;
		if (blockIdx.x == 0)
			global_box[k] = v;
	}
	__syncthreads();
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const f3 lo = mk3(gb[0], gb[1], gb[2]);
	const f3 hi = mk3(gb[3], gb[4], gb[5]);
	const f3 ext = max3(hi - lo, mk3(1e-6f, 1e-6f, 1e-6f));
	const f3 inv = mk3(1.0f / ext.x, 1.0f / ext.y, 1.0f / ext.z);
	const f3 nrm = mul3(ld3(center, i) - lo, inv);
	keys[i] = morton30(nrm);
	vals[i] = i;
}

// ---------------------------------------------------------------------------
// Karras build (lbvh.cpp:206-293, 530-600).  Keys are (morton30 << 32 | orig).
__device__ __forceinline__ uint64_t skey(const uint32_t* __restrict__ k, const uint32_t* __restrict__ v, int i)
{
	return ((uint64_t)k[i] << 32) | (uint64_t)v[i];
}
__device__ __forceinline__ int common_prefix(uint64_t a, uint64_t b)
{
	const uint64_t x = a ^ b;
	return x ? __clzll((long long)x) : 64;
}
__device__ __forceinline__ int cp_ij(uint64_t mi, int j, const uint32_t* k, const uint32_t* v, int n)
{
	return (j >= 0 && j < n) ? common_prefix(mi, skey(k, v, j)) : -1;
}

__global__ void k_karras(uint32_t n,
	const uint32_t* __restrict__ skeys,
	const uint32_t* __restrict__ svals,
	uint64_t* __restrict__ morton_sorted,
	uint32_t* __restrict__ sorted_get_original,
	uint32_t* __restrict__ parent,
	uint32_t* __restrict__ children,
	uint32_t* __restrict__ object_idx,
	BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ leaf_ps,
	uint32_t* __restrict__ inner_ps,
	uint32_t* __restrict__ refit_flag,
	uint32_t* __restrict__ status)
{
	const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	const int nl = (int)n;
	const int ni = nl - 1;
	if (i >= nl)
		return;
	// fn_apply_sorted
	const uint32_t orig = svals[i];
	morton_sorted[i] = skey(skeys, svals, i);
	sorted_get_original[i] = orig;
	children[2 * (ni + i) + 0] = orig;
	children[2 * (ni + i) + 1] = orig;
	object_idx[ni + i] = orig;
	if (i == 0)
	{
		status[0] = 1u;
		status[1] = 0u;
		parent[0] = kNone;
	}
	if (ni == 0)
	{
		// Single primitive: a stand-in root whose second slot never overlaps.
		BvhNode r;
		r.a = make_float4(INFINITY, INFINITY, INFINITY, -INFINITY);
		r.b = make_float4(-INFINITY, -INFINITY, INFINITY, INFINITY);
		r.c = make_float4(INFINITY, -INFINITY, -INFINITY, -INFINITY);
		r.d = make_uint4(kLeafBit | orig, kLeafBit | orig, 0u, 0u);
		nodes[0] = r;
		leaf_ps[0] = 2u; // node 0, local, slot 0
		return;
	}
	if (i >= ni)
		return;
	// fn_build_inner_nodes
	int r0 = 0, r1 = ni;
	const uint64_t mi0 = skey(skeys, svals, i);
	if (i != 0)
	{
		const int cl = cp_ij(mi0, i - 1, skeys, svals, nl);
		const int cr = cp_ij(mi0, i + 1, skeys, svals, nl);
		const int d = cl < cr ? 1 : -1;
		const int cmin = min(cl, cr);
		int		  lmax = 2;
		while (cp_ij(mi0, i + lmax * d, skeys, svals, nl) > cmin)
			lmax <<= 1;
		int l = 0;
		for (int t = lmax >> 1; t >= 1; t >>= 1)
			if (cp_ij(mi0, i + (l + t) * d, skeys, svals, nl) > cmin)
				l += t;
		r0 = i;
		r1 = i + l * d;
	}
	// findSplit
	int a = r0, b = r1;
	if (a > b)
	{
		int t = a;
		a = b;
		b = t;
	}
	const uint64_t ma = skey(skeys, svals, a);
	const uint64_t mb = skey(skeys, svals, b);
	int			   split;
	if (ma == mb)
		split = (a + b) >> 1;
	else
	{
		const int cpn = common_prefix(ma, mb);
		int		  t = b - a;
		split = a;
		while (true)
		{
			t = (t + 1) >> 1;
			const int ns = split + t;
			if (ns < b && common_prefix(ma, skey(skeys, svals, ns)) > cpn)
				split = ns;
			if (!(t > 1))
				break;
		}
	}
	int cl = (min(r0, r1) == split) ? (ni + split) : split;
	int cr = (max(r0, r1) == split + 1) ? (ni + split + 1) : (split + 1);
	if (cr >= ni)
	{
		int t = cl;
		cl = cr;
		cr = t;
	}
	parent[cl] = (uint32_t)i;
	parent[cr] = (uint32_t)i;
	children[2 * i + 0] = (uint32_t)cl;
	children[2 * i + 1] = (uint32_t)cr;
	object_idx[i] = kNone;
	// traversal record + refit bookkeeping: children learn parent, slot and
	// whether the parent's leaf range [min(r0,r1), max(r0,r1)] lies in one refit block
	const uint32_t ref0 = cl >= ni ? (kLeafBit | svals[cl - ni]) : (uint32_t)cl;
	const uint32_t ref1 = cr >= ni ? (kLeafBit | svals[cr - ni]) : (uint32_t)cr;
	nodes[i].d = make_uint4(ref0, ref1, 0u, 0u);
	const uint32_t local = ((uint32_t)min(r0, r1) / kRefitBlock == (uint32_t)max(r0, r1) / kRefitBlock) ? 2u : 0u;
	const uint32_t ps0 = ((uint32_t)i << 2) | local | 0u;
	const uint32_t ps1 = ((uint32_t)i << 2) | local | 1u;
	if (cl >= ni)
		leaf_ps[cl - ni] = ps0;
	else
		inner_ps[cl] = ps0;
	if (cr >= ni)
		leaf_ps[cr - ni] = ps1;
	else
		inner_ps[cr] = ps1;
	if (i == 0)
		inner_ps[0] = kNone;
	refit_flag[i] = 0u;
}

// ---------------------------------------------------------------------------
// Leaf swept box + thickness (fn_update_*_leave_aabb_v2, lbvh.cpp:621-688),
// then bottom-up refit where the second child to arrive at a node builds the
// union (replaces the CAS/spin-wait scheme of lbvh.cpp:694-758; no thread ever
// waits on another).
__device__ __forceinline__ void store_slot(BvhNode* node, uint32_t slot, const box3& b)
{
	float* f = reinterpret_cast<float*>(node) + 6 * slot;
	f[0] = b.lo.x;
	f[1] = b.lo.y;
	f[2] = b.lo.z;
	f[3] = b.hi.x;
	f[4] = b.hi.y;
	f[5] = b.hi.z;
}
__device__ __forceinline__ box3 load_slot(const BvhNode* node, uint32_t slot)
{
	const float* f = reinterpret_cast<const float*>(node) + 6 * slot;
	return box3{ mk3(f[0], f[1], f[2]), mk3(f[3], f[4], f[5]) };
}

__device__ __forceinline__ void store_root(float* root_box, const box3& b)
{
	root_box[0] = b.lo.x;
	root_box[1] = b.lo.y;
	root_box[2] = b.lo.z;
	root_box[3] = b.hi.x;
	root_box[4] = b.hi.y;
	root_box[5] = b.hi.z;
}

// Phase 1: one thread per sorted leaf.  Climbs while the parent is local to
// this workgroup (both children are then refit by threads of the same
// workgroup, so a workgroup-scope acq_rel counter orders the box writes).  At
// the first non-local parent the thread stores its box into the parent's slot
// and records the arrival in `exits` for phase 2.
__global__ __launch_bounds__(kRefitBlock) void k_refit_local(int type,
	uint32_t n,
	const uint32_t* __restrict__ sorted_get_original,
	const uint32_t* __restrict__ prims,
	const float* __restrict__ thickness,
	const float* __restrict__ xs,
	const float* __restrict__ xe,
	BvhNode* nodes,
	const uint32_t* __restrict__ leaf_ps,
	const uint32_t* __restrict__ inner_ps,
	uint32_t* refit_flag,
	uint32_t* __restrict__ exits,
	uint32_t* __restrict__ exit_count,
	float* __restrict__ root_box,
	uint32_t* __restrict__ status)
{
	const uint32_t lid = blockIdx.x * blockDim.x + threadIdx.x;
	if (lid >= n)
		return;
	const uint32_t orig = sorted_get_original[lid];
	f3			   ps[3], pe[3];
	uint32_t	   v0;
	prim_points(type, xs, prims, orig, ps, &v0);
	prim_points(type, xe, prims, orig, pe, &v0);
	box3 bx;
	if (type == kTreeFace)
	{
		box3 s = box_add(box_add(box_of(ps[0]), ps[1]), ps[2]);
		box3 e = box_add(box_add(box_of(pe[0]), pe[1]), pe[2]);
		bx = box_union(s, e);
	}
	else
		bx = box_add(box_add(box_add(box_of(ps[0]), ps[1]), pe[0]), pe[1]);
	bx = box_inflate(bx, thickness[v0]);

	uint32_t slot_ref = leaf_ps[lid];
	// A tree from unique 64-bit keys is at most 64 levels deep; the cap only
	// guards against a corrupted tree (reported through status[0]).
	for (int level = 0;; ++level)
	{
		if (level > 96)
		{
			status[0] = 0u;
			return;
		}
		const uint32_t node = slot_ref >> 2;
		store_slot(&nodes[node], slot_ref & 1u, bx);
		if (n == 1)
		{
			store_root(root_box, bx);
			return;
		}
		if (!(slot_ref & 2u))
		{
			exits[atomicAdd(exit_count, 1u)] = node;
			return;
		}
		const uint32_t old =
			__hip_atomic_fetch_add(&refit_flag[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (old == 0u)
			return; // sibling not done yet; it carries the union upward
		refit_flag[node] = 0u;
		bx = box_union(load_slot(&nodes[node], 0), load_slot(&nodes[node], 1));
		slot_ref = inner_ps[node];
		if (slot_ref == kNone)
		{
			store_root(root_box, bx);
			return;
		}
	}
}

// Phase 2: one workgroup finishes the nodes whose leaf range spans several
// refit blocks (about two per block).  Phase-1 boxes are visible through the
// kernel boundary; arrivals inside this workgroup use workgroup scope.
__global__ __launch_bounds__(1024) void k_refit_top(BvhNode* nodes,
	const uint32_t* __restrict__ inner_ps,
	uint32_t* refit_flag,
	const uint32_t* __restrict__ exits,
	uint32_t* __restrict__ exit_count,
	float* __restrict__ root_box,
	uint32_t* __restrict__ status)
{
	const uint32_t ne = *exit_count;
	for (uint32_t e = threadIdx.x; e < ne; e += blockDim.x)
	{
		uint32_t node = exits[e];
		for (int level = 0;; ++level)
		{
			if (level > 96)
			{
				status[0] = 0u;
				break;
			}
			const uint32_t old =
				__hip_atomic_fetch_add(&refit_flag[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
			if (old == 0u)
				break;
			refit_flag[node] = 0u;
			const box3	   bx = box_union(load_slot(&nodes[node], 0), load_slot(&nodes[node], 1));
			const uint32_t ps = inner_ps[node];
			if (ps == kNone)
			{
				store_root(root_box, bx);
				break;
			}
			node = ps >> 2;
			store_slot(&nodes[node], ps & 1u, bx);
		}
	}
	__syncthreads();
	if (threadIdx.x == 0)
		*exit_count = 0u;
}

// ---------------------------------------------------------------------------
// Broad-phase queries (query_template2, lbvh.cpp:295-363) with staged output.
__global__ __launch_bounds__(kTravThreads) void k_query(int kind,
	uint32_t nq,
	const float* __restrict__ xb,
	const float* __restrict__ xe,
	const uint32_t* __restrict__ edges,
	const float* __restrict__ d_hat,
	const float* __restrict__ thickness,
	const BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ count,
	uint2* __restrict__ list,
	uint32_t max_count,
	uint32_t* __restrict__ status,
	uint32_t* __restrict__ spill,
	uint32_t max_visits)
{
	__shared__ TravShared sh;
	const uint32_t		  q = blockIdx.x * kTravThreads + threadIdx.x;
	const bool			  active = q < nq;
	const box3			  qb = active ? query_box(kind, q, xb, xe, edges, d_hat, thickness) : box_empty();
	EmitPairs			  emit{ count, list, max_count };
	traverse_staged(sh, active, q, qb, kind == kQueryEdges, nodes, status, spill, max_visits, emit);
}

// ---------------------------------------------------------------------------
void Lbvh::init(Context* c, int tree_type, uint32_t n)
{
	UT_REQUIRE(tree_type == kTreeFace || tree_type == kTreeEdge, "tree type must be face (1) or edge (2)");
	UT_REQUIRE(n >= 1 && n < (1u << 30), "LBVH needs 1 <= num_leaves < 2^30");
	ctx = c;
	type = tree_type;
	num_leaves = n;
	const uint32_t nodes_n = 2 * n - 1;
	const uint32_t inner = n > 1 ? n - 1 : 1;
	leaf_center.reserve(3 * (size_t)n);
	block_box.reserve(6 * (size_t)div_up(n, kBoxThreads));
	global_box.reserve(6);
	for (int k = 0; k < 2; ++k)
	{
		sort_keys[k].reserve(n);
		sort_vals[k].reserve(n);
		sort.keys[k] = sort_keys[k].get();
		sort.vals[k] = sort_vals[k].get();
	}
	sort_hist.reserve(radix_sort_hist_size(n));
	sort.hist = sort_hist.get();
	morton_sorted.reserve(n);
	sorted_get_original.reserve(n);
	parent.reserve(nodes_n);
	children.reserve(2 * (size_t)nodes_n);
	object_idx.reserve(nodes_n);
	nodes.reserve(inner);
	leaf_parent_slot.reserve(n);
	inner_parent_slot.reserve(inner);
	refit_flag.reserve(inner);
	refit_exits.reserve(2 * (size_t)n + 2);
	refit_exit_count.reserve(1);
	root_box.reserve(6);
	status.reserve(2);
	ctx->activate();
	UT_HIP_CHECK(hipMemsetAsync(refit_exit_count.get(), 0, sizeof(uint32_t), ctx->stream));
	UT_HIP_CHECK(hipMemsetAsync(refit_flag.get(), 0, sizeof(uint32_t) * inner, ctx->stream));
	UT_HIP_CHECK(hipMemsetAsync(status.get(), 0, sizeof(uint32_t) * 2, ctx->stream));
}

void Lbvh::reduce_aabb(const float* x, const uint32_t* prims)
{
	UT_REQUIRE(x && prims, "reduce_aabb: null input");
	const uint32_t blocks = div_up(num_leaves, kBoxThreads);
	k_leaf_center_box<<<blocks, kBoxThreads, 0, ctx->stream>>>(
		type, x, prims, num_leaves, leaf_center.get(), block_box.get());
	k_morton<<<div_up(num_leaves, 256), 256, 0, ctx->stream>>>(num_leaves,
		leaf_center.get(),
		block_box.get(),
		blocks,
		global_box.get(),
		sort.keys[0],
		sort.vals[0]);
	UT_LAUNCH_CHECK();
}

void Lbvh::construct_tree()
{
	sort.result = 0;
	radix_sort_pairs(sort, num_leaves, 30, ctx->stream);
	construct_tree_presorted();
}

void Lbvh::construct_tree_presorted()
{
	k_karras<<<div_up(num_leaves, 256), 256, 0, ctx->stream>>>(num_leaves,
		sort.keys[sort.result],
		sort.vals[sort.result],
		morton_sorted.get(),
		sorted_get_original.get(),
		parent.get(),
		children.get(),
		object_idx.get(),
		nodes.get(),
		leaf_parent_slot.get(),
		inner_parent_slot.get(),
		refit_flag.get(),
		status.get());
	UT_LAUNCH_CHECK();
}

void Lbvh::update_leaves_and_refit(const float* thickness, const float* xs, const float* xe, const uint32_t* prims)
{
	UT_REQUIRE(thickness && xs && xe && prims, "update_leaves_and_refit: null input");
	k_refit_local<<<div_up(num_leaves, kRefitBlock), kRefitBlock, 0, ctx->stream>>>(type,
		num_leaves,
		sorted_get_original.get(),
		prims,
		thickness,
		xs,
		xe,
		nodes.get(),
		leaf_parent_slot.get(),
		inner_parent_slot.get(),
		refit_flag.get(),
		refit_exits.get(),
		refit_exit_count.get(),
		root_box.get(),
		status.get());
	k_refit_top<<<1, 1024, 0, ctx->stream>>>(nodes.get(),
		inner_parent_slot.get(),
		refit_flag.get(),
		refit_exits.get(),
		refit_exit_count.get(),
		root_box.get(),
		status.get());
	UT_LAUNCH_CHECK();
}

void Lbvh::query(int kind,
	const float*	xb,
	const float*	xe,
	const uint32_t* edges,
	uint32_t		nq,
	const float*	d_hat,
	const float*	thickness,
	uint32_t*		count,
	uint32_t*		list,
	uint32_t		max_count)
{
	UT_REQUIRE(kind >= kQueryVerts && kind <= kQueryEdgesUnculled, "unknown query kind");
	UT_REQUIRE(xb && xe && d_hat && thickness && count, "query: null input");
	UT_REQUIRE(kind == kQueryVerts || edges, "edge query needs edges");
	UT_REQUIRE(max_count == 0 || list, "query: null output list");
	if (nq == 0)
		return;
	spill.reserve((size_t)nq * (kStackTotal - kStackLds));
	k_query<<<div_up(nq, kTravThreads), kTravThreads, 0, ctx->stream>>>(kind,
		nq,
		xb,
		xe,
		edges,
		d_hat,
		thickness,
		nodes.get(),
		count,
		reinterpret_cast<uint2*>(list),
		max_count,
		status.get(),
		spill.get(),
		num_leaves + 1);
	UT_LAUNCH_CHECK();
}

} // namespace ut
