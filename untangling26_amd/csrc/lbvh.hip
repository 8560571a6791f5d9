// LBVH kernels for gfx950.  See lbvh.hpp for the data layout.
#include "ut_common.hpp"
#include "lbvh.hpp"

namespace ut
{

constexpr int kBoxThreads = 256;
constexpr int kQueryThreads = 256;
constexpr int kStack = 32;

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

// Leaf centroid + block box, last block folds all block boxes into the
// global box (lbvh.cpp:400-496 do this in two launches).
__global__ __launch_bounds__(kBoxThreads) void k_leaf_center_box(int type,
	const float* __restrict__ x,
	const uint32_t* __restrict__ prims,
	uint32_t n,
	float* __restrict__ center,
	float* __restrict__ block_box,
	uint32_t* __restrict__ done,
	float* __restrict__ global_box)
{
	__shared__ float red[6][kBoxThreads / kWave];
	__shared__ bool	 is_last;
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
	if (threadIdx.x == 0)
	{
		float r[6];
#pragma unroll
		for (int k = 0; k < 6; ++k)
			r[k] = red[k][0];
		for (int q = 1; q < kBoxThreads / kWave; ++q)
		{
#pragma unroll
			for (int k = 0; k < 3; ++k)
				r[k] = fminf(r[k], red[k][q]);
#pragma unroll
			for (int k = 3; k < 6; ++k)
				r[k] = fmaxf(r[k], red[k][q]);
		}
#pragma unroll
		for (int k = 0; k < 6; ++k)
			block_box[6 * blockIdx.x + k] = r[k];
		const uint32_t ticket =
			__hip_atomic_fetch_add(done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		is_last = (ticket == gridDim.x - 1);
	}
	__syncthreads();
	if (!is_last)
		return;
	__atomic_thread_fence(__ATOMIC_ACQUIRE); // other blocks' boxes are visible
	float r[6] = { 1e8f, 1e8f, 1e8f, -1e8f, -1e8f, -1e8f };
	for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBoxThreads)
	{
#pragma unroll
		for (int k = 0; k < 3; ++k)
			r[k] = fminf(r[k], __builtin_nontemporal_load(&block_box[6 * b + k]));
#pragma unroll
		for (int k = 3; k < 6; ++k)
			r[k] = fmaxf(r[k], __builtin_nontemporal_load(&block_box[6 * b + k]));
	}
#pragma unroll
	for (int k = 0; k < 3; ++k)
		r[k] = wave_min(r[k]);
#pragma unroll
	for (int k = 3; k < 6; ++k)
		r[k] = wave_max(r[k]);
	__syncthreads();
	if (__lane_id() == 0)
	{
#pragma unroll
		for (int k = 0; k < 6; ++k)
			red[k][w] = r[k];
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (int q = 1; q < kBoxThreads / kWave; ++q)
		{
#pragma unroll
			for (int k = 0; k < 3; ++k)
				red[k][0] = fminf(red[k][0], red[k][q]);
#pragma unroll
			for (int k = 3; k < 6; ++k)
				red[k][0] = fmaxf(red[k][0], red[k][q]);
		}
#pragma unroll
		for (int k = 0; k < 6; ++k)
			global_box[k] = red[k][0];
		*done = 0u; // ready for the next build
	}
}

// fn_compute_mortons, lbvh.cpp:512-528.
__global__ void k_morton(uint32_t n,
	const float* __restrict__ center,
	const float* __restrict__ global_box,
	uint32_t* __restrict__ keys,
	uint32_t* __restrict__ vals)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const f3 lo = mk3(global_box[0], global_box[1], global_box[2]);
	const f3 hi = mk3(global_box[3], global_box[4], global_box[5]);
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
		leaf_ps[0] = 0u;
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
	// traversal record + refit bookkeeping
	const uint32_t ref0 = cl >= ni ? (kLeafBit | svals[cl - ni]) : (uint32_t)cl;
	const uint32_t ref1 = cr >= ni ? (kLeafBit | svals[cr - ni]) : (uint32_t)cr;
	nodes[i].d = make_uint4(ref0, ref1, 0u, 0u);
	if (cl >= ni)
		leaf_ps[cl - ni] = ((uint32_t)i << 1) | 0u;
	else
		inner_ps[cl] = ((uint32_t)i << 1) | 0u;
	if (cr >= ni)
		leaf_ps[cr - ni] = ((uint32_t)i << 1) | 1u;
	else
		inner_ps[cr] = ((uint32_t)i << 1) | 1u;
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

__global__ void k_leaf_refit(int type,
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
		const uint32_t node = slot_ref >> 1;
		store_slot(&nodes[node], slot_ref & 1u, bx);
		if (n == 1)
		{
			for (int k = 0; k < 3; ++k)
				root_box[k] = (&bx.lo.x)[k], root_box[3 + k] = (&bx.hi.x)[k];
			return;
		}
		const uint32_t old =
			__hip_atomic_fetch_add(&refit_flag[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		if (old == 0u)
			return; // sibling not done yet; it will carry the union upward
		refit_flag[node] = 0u;
		const box3 b0 = load_slot(&nodes[node], 0);
		const box3 b1 = load_slot(&nodes[node], 1);
		bx = box_union(b0, b1);
		slot_ref = inner_ps[node];
		if (slot_ref == kNone)
		{
			root_box[0] = bx.lo.x;
			root_box[1] = bx.lo.y;
			root_box[2] = bx.lo.z;
			root_box[3] = bx.hi.x;
			root_box[4] = bx.hi.y;
			root_box[5] = bx.hi.z;
			return;
		}
	}
}

// ---------------------------------------------------------------------------
// Stack traversal (query_template2, lbvh.cpp:295-363), one thread per query
// primitive.  The first kStack entries of the stack live in LDS; deeper
// entries spill to a per-query global area.  Keys are unique 64-bit values, so
// the Karras tree is at most 64 levels deep and kStackTotal = 128 is never
// reached; if it were, status[1] counts the truncated queries (the reference
// silently drops them, lbvh.cpp:343-351).
constexpr int kStackTotal = 128;

__global__ __launch_bounds__(kQueryThreads) void k_query(int kind,
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
	__shared__ uint32_t lds_stack[kStack * kQueryThreads];
	const uint32_t		q = blockIdx.x * kQueryThreads + threadIdx.x;
	const bool			active = q < nq;
	box3				qb = box_empty();
	if (active)
	{
		if (kind == kQueryVerts)
		{
			qb = box_add(box_of(ld3(xb, q)), ld3(xe, q));
			qb = box_inflate(qb, thickness[q] + d_hat[q]);
		}
		else
		{
			const uint32_t e0 = edges[2 * q], e1 = edges[2 * q + 1];
			qb = box_add(box_add(box_add(box_of(ld3(xb, e0)), ld3(xb, e1)), ld3(xe, e0)), ld3(xe, e1));
			qb = box_inflate(qb, thickness[e0] + d_hat[e0]);
		}
	}
	int		 sp = 0;
	uint32_t node = 0;
	bool	 go = active;
	uint32_t visits = 0;
	while (__ballot(go) != 0ull)
	{
		bool	 emit0 = false, emit1 = false;
		uint32_t obj0 = 0, obj1 = 0;
		if (go && ++visits > max_visits)
		{
			// every inner node is visited at most once per query; more means a corrupted tree
			status[0] = 0u;
			go = false;
		}
		if (go)
		{
			const BvhNode* rec = &nodes[node];
			const float4   a = rec->a, b = rec->b, c = rec->c;
			const uint2	   refs = make_uint2(rec->d.x, rec->d.y);
			const box3	   b0{ mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y) };
			const box3	   b1{ mk3(b.z, b.w, c.x), mk3(c.y, c.z, c.w) };
			const bool	   h0 = box_overlap(b0, qb);
			const bool	   h1 = box_overlap(b1, qb);
			const bool	   l0 = (refs.x & kLeafBit) != 0u, l1 = (refs.y & kLeafBit) != 0u;
			obj0 = refs.x & ~kLeafBit;
			obj1 = refs.y & ~kLeafBit;
			emit0 = h0 && l0 && (kind != kQueryEdges || q < obj0);
			emit1 = h1 && l1 && (kind != kQueryEdges || q < obj1);
			const bool t0 = h0 && !l0, t1 = h1 && !l1;
			if (t0 && t1)
			{
				if (sp < kStack)
					lds_stack[sp * kQueryThreads + threadIdx.x] = refs.y;
				else if (sp < kStackTotal)
					spill[(size_t)q * (kStackTotal - kStack) + (sp - kStack)] = refs.y;
				if (sp < kStackTotal)
				{
					++sp;
					node = refs.x;
				}
				else
				{
					atomicAdd(&status[1], 1u);
					go = false;
				}
			}
			else if (t0)
				node = refs.x;
			else if (t1)
				node = refs.y;
			else if (sp > 0)
			{
				--sp;
				node = sp < kStack ? lds_stack[sp * kQueryThreads + threadIdx.x]
								   : spill[(size_t)q * (kStackTotal - kStack) + (sp - kStack)];
			}
			else
				go = false;
		}
		const uint32_t s0 = wave_append(count, emit0);
		if (emit0 && s0 < max_count)
			list[s0] = make_uint2(q, obj0);
		const uint32_t s1 = wave_append(count, emit1);
		if (emit1 && s1 < max_count)
			list[s1] = make_uint2(q, obj1);
	}
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
	done_counter.reserve(1);
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
	root_box.reserve(6);
	status.reserve(2);
	ctx->activate();
	UT_HIP_CHECK(hipMemsetAsync(done_counter.get(), 0, sizeof(uint32_t), ctx->stream));
	UT_HIP_CHECK(hipMemsetAsync(refit_flag.get(), 0, sizeof(uint32_t) * inner, ctx->stream));
	UT_HIP_CHECK(hipMemsetAsync(status.get(), 0, sizeof(uint32_t) * 2, ctx->stream));
}

void Lbvh::reduce_aabb(const float* x, const uint32_t* prims)
{
	UT_REQUIRE(x && prims, "reduce_aabb: null input");
	const uint32_t blocks = div_up(num_leaves, kBoxThreads);
	k_leaf_center_box<<<blocks, kBoxThreads, 0, ctx->stream>>>(
		type, x, prims, num_leaves, leaf_center.get(), block_box.get(), done_counter.get(), global_box.get());
	k_morton<<<div_up(num_leaves, 256), 256, 0, ctx->stream>>>(
		num_leaves, leaf_center.get(), global_box.get(), sort.keys[0], sort.vals[0]);
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
	k_leaf_refit<<<div_up(num_leaves, 256), 256, 0, ctx->stream>>>(type,
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
	spill.reserve((size_t)nq * (kStackTotal - kStack));
	k_query<<<div_up(nq, kQueryThreads), kQueryThreads, 0, ctx->stream>>>(kind,
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
