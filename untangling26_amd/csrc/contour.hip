// Intersection contours on gfx950: canonical EF order, EF adjacency, loop
// vertices, and connected components with mesh parity.
//
// Components and parity come from one lock-free union-find over the
// "parity-doubled" graph: every EF pair p owns two nodes 2p (p on side 0) and
// 2p+1 (p on side 1); an adjacency p-q with flip weight w joins 2p with
// 2q+w and 2p+1 with 2q+1-w.  Roots always hook under the smaller root, so a
// set's root is its smallest node.  For a component whose smallest pair is m,
// node 2m is the root of one of the two sets p's nodes belong to:
//   comp_min(p) = min(find(2p), find(2p+1)) / 2
//   parity(p)   = 0 iff find(2p) == 2m      (p on the same side as m)
//   odd cycle   iff find(2p) == find(2p+1)
// This equals the DFS parity of contour_flood_filling_template (which starts
// each component at its smallest pair with parity 0) on every consistent
// component, and the component numbering (rank of the smallest pair) equals
// the reference's visiting order.
#include "ut_common.hpp"
#include "contour.hpp"
#include "radix_sort.hpp"
#include "scan.hpp"

namespace ut
{

__device__ __forceinline__ uint32_t ld_relaxed(uint32_t* p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// parent[x] <= x always holds, so the walk strictly decreases and ends.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x)
{
	while (true)
	{
		const uint32_t p = ld_relaxed(&parent[x]);
		if (p == x)
			return x;
		const uint32_t gp = ld_relaxed(&parent[p]);
		if (gp != p)
			st_relaxed(&parent[x], gp); // path halving; gp is an ancestor of x
		x = gp;
	}
}

// Every failed CAS means another thread hooked a root, which happens at most
// n-1 times in total, so the loop ends.
__device__ __forceinline__ void uf_union(uint32_t* parent, uint32_t a, uint32_t b)
{
	uint32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
	while (ra != rb)
	{
		if (ra < rb)
		{
			const uint32_t t = ra;
			ra = rb;
			rb = t;
		}
		const uint32_t old = atomicCAS(&parent[ra], ra, rb);
		if (old == ra)
			return;
		ra = uf_find(parent, old);
		rb = uf_find(parent, rb);
	}
}

__global__ void k_uf_init(uint32_t* parent, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		parent[i] = i;
}

__global__ void k_uf_compress(uint32_t* parent, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		st_relaxed(&parent[i], uf_find(parent, i));
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

// comp_min / parity / root flag / odd-cycle flag from the doubled-graph roots.
__global__ void k_uf_labels(const uint32_t* __restrict__ parent, uint32_t n, uint32_t* __restrict__ comp_min,
	uint32_t* __restrict__ parity, uint32_t* __restrict__ is_root, uint32_t* __restrict__ inconsistent)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n)
		return;
	const uint32_t r0 = parent[2 * p], r1 = parent[2 * p + 1];
	const uint32_t m2 = min(r0, r1);
	const uint32_t m = m2 >> 1;
	comp_min[p] = m;
	parity[p] = (r0 == m2) ? 0u : 1u;
	if (is_root)
		is_root[p] = (m == p) ? 1u : 0u;
	if (r0 == r1)
		atomicOr(&inconsistent[m], 1u);
}

// ---- CSR graph interface -------------------------------------------------
__global__ void k_graph_union(uint32_t n, const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
	const uint32_t* __restrict__ w, uint32_t* parent)
{
	const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
	if (u >= n)
		return;
	for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e)
	{
		const uint32_t v = col[e], f = w[e] & 1u;
		uf_union(parent, 2 * u, 2 * v + f);
		uf_union(parent, 2 * u + 1, 2 * v + (1u - f));
	}
}

void graph_components(Context* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
	const uint32_t* flip_w, uint32_t* uf, uint32_t* comp_min, uint32_t* parity, uint32_t* inconsistent,
	uint32_t* status)
{
	(void)status;
	if (n == 0)
		return;
	hipStream_t s = ctx->stream;
	k_uf_init<<<div_up(2 * n, 256), 256, 0, s>>>(uf, 2 * n);
	k_fill_u32<<<div_up(n, 256), 256, 0, s>>>(inconsistent, n, 0u);
	k_graph_union<<<div_up(n, 256), 256, 0, s>>>(n, row_ptr, col_idx, flip_w, uf);
	k_uf_compress<<<div_up(2 * n, 256), 256, 0, s>>>(uf, 2 * n);
	k_uf_labels<<<div_up(n, 256), 256, 0, s>>>(uf, n, comp_min, parity, nullptr, inconsistent);
	UT_LAUNCH_CHECK();
}

// ---- EF pairs ---------------------------------------------------------------
__global__ void k_ct_key_fid(uint32_t P, const uint32_t* __restrict__ ids, uint32_t* __restrict__ k,
	uint32_t* __restrict__ v)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P)
		return;
	k[i] = ids[2 * i + 1];
	v[i] = i;
}
__global__ void k_ct_key_eid(uint32_t P, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ perm,
	uint32_t* __restrict__ k)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P)
		return;
	k[i] = ids[2 * perm[i]];
}

// Gather into (edge, face) order (intersection_resolver.cpp:695-726).
__global__ void k_ct_gather(uint32_t P, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ in_ids,
	const EfPairRec* __restrict__ in_pairs, uint32_t* __restrict__ ids, EfPairRec* __restrict__ pairs,
	uint64_t* __restrict__ keys)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P)
		return;
	const uint32_t j = perm[i];
	const uint32_t e = in_ids[2 * j], f = in_ids[2 * j + 1];
	ids[2 * i] = e;
	ids[2 * i + 1] = f;
	pairs[i] = in_pairs[j];
	keys[i] = ((uint64_t)e << 32) | f;
}

// memcmp order of two 48-byte records: the first differing byte decides, as
// unsigned bytes in memory order (little-endian words: lowest byte first).
__device__ __forceinline__ int payload_cmp(const EfPairRec& a, const EfPairRec& b)
{
	const uint32_t* x = reinterpret_cast<const uint32_t*>(&a);
	const uint32_t* y = reinterpret_cast<const uint32_t*>(&b);
	for (int w = 0; w < 12; ++w)
	{
		const uint32_t d = x[w] ^ y[w];
		if (d)
		{
			const int sh = __builtin_ctz(d) & ~7;
			return ((x[w] >> sh) & 0xFFu) < ((y[w] >> sh) & 0xFFu) ? -1 : 1;
		}
	}
	return 0;
}

// Pairs with equal (edge, face) keys (e.g. a DCD pair and pairs made from
// zero-distance proximity pairs) ordered by payload bytes, the tie-break of
// intersection_resolver.cpp:712-716.  One thread per run of equal keys.
__global__ void k_ct_tie_order(uint32_t P, const uint64_t* __restrict__ keys, EfPairRec* __restrict__ pairs)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P || (i > 0 && keys[i - 1] == keys[i]) || i + 1 >= P || keys[i + 1] != keys[i])
		return;
	uint32_t end = i + 1;
	while (end < P && keys[end] == keys[i])
		++end;
	for (uint32_t a = i + 1; a < end; ++a) // insertion sort; runs are a few pairs long
	{
		const EfPairRec r = pairs[a];
		uint32_t		b = a;
		while (b > i && payload_cmp(pairs[b - 1], r) > 0)
		{
			pairs[b] = pairs[b - 1];
			--b;
		}
		pairs[b] = r;
	}
}

__device__ __forceinline__ bool face_has(const uint32_t* faces, uint32_t f, uint32_t v)
{
	return faces[3 * f] == v || faces[3 * f + 1] == v || faces[3 * f + 2] == v;
}

// find_loop_vertex (intersection_resolver.cpp:431-501): the vertex of a face
// adjacent to the edge that lies opposite the edge and is a corner of the
// intersected face.  Returns the loop edge or kNone.
__device__ __forceinline__ uint32_t loop_vertex(uint32_t e, uint32_t f, uint32_t T, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, uint32_t* vid)
{
	const uint32_t a = edges[2 * e], b = edges[2 * e + 1];
	const uint32_t v = faces[3 * T] + faces[3 * T + 1] + faces[3 * T + 2] - a - b;
	if (!face_has(faces, f, v))
		return kNone;
	*vid = v;
	const uint32_t fae0 = face_adj_edges[3 * T], fae1 = face_adj_edges[3 * T + 1], fae2 = face_adj_edges[3 * T + 2];
	auto has = [&](uint32_t ed) { return edges[2 * ed] == v || edges[2 * ed + 1] == v; };
	if (e == fae0)
		return has(fae1) ? fae1 : fae2;
	if (e == fae1)
		return has(fae0) ? fae0 : fae2;
	if (e == fae2)
		return has(fae0) ? fae0 : fae1;
	return kNone;
}

__device__ __forceinline__ uint32_t lower_key(const uint64_t* __restrict__ keys, uint32_t P, uint64_t k)
{
	uint32_t lo = 0, hi = P;
	while (lo < hi)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (keys[mid] < k)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

// compute_adj_pairs (intersection_resolver.cpp:503-558) of pair p among the P
// sorted pairs, in the reference's order: every pair (all of a key's pairs)
// on another edge of a face around e with the same intersected face, then
// every pair of an edge of the intersected face cutting a face around e.
// visit(q) returns false to stop.
template <class Visit>
__device__ __forceinline__ void for_each_adjacent(uint32_t p, uint32_t P, const uint32_t* __restrict__ ids,
	const uint64_t* __restrict__ keys, const uint32_t* __restrict__ face_adj_edges,
	const uint32_t* __restrict__ edge_adj_faces, Visit&& visit)
{
	const uint32_t e = ids[2 * p], f = ids[2 * p + 1];
	const uint32_t T[2] = { edge_adj_faces[2 * e], edge_adj_faces[2 * e + 1] };
	auto		   key_run = [&](uint32_t ke, uint32_t kf) -> bool
	{
		const uint64_t k = ((uint64_t)ke << 32) | kf;
		for (uint32_t q = lower_key(keys, P, k); q < P && keys[q] == k; ++q)
			if (q != p && !visit(q))
				return false;
		return true;
	};
	for (int k = 0; k < 2; ++k)
	{
		if (T[k] == kNone)
			continue;
		const uint32_t a = face_adj_edges[3 * T[k]], b = face_adj_edges[3 * T[k] + 1], c = face_adj_edges[3 * T[k] + 2];
		const uint32_t o0 = (a == e) ? b : a;
		const uint32_t o1 = (a == e || b == e) ? c : b;
		if (!key_run(o0, f) || !key_run(o1, f))
			return;
	}
	for (int ii = 0; ii < 3; ++ii)
	{
		const uint32_t fe = face_adj_edges[3 * f + ii];
		if (fe == e)
			continue;
		for (int k = 0; k < 2; ++k)
			if (T[k] != kNone && !key_run(fe, T[k]))
				return;
	}
}

// is_degenerate_ef_pair (intersection_resolver.cpp:561-569).
__device__ __forceinline__ bool ef_degenerate(const EfPairRec& r)
{
	const float e1 = 1.0f - r.bary[0];
	const float f2 = 1.0f - r.bary[1] - r.bary[2];
	const float emin = e1 < r.bary[0] ? e1 : r.bary[0];
	float		fmin = r.bary[1];
	if (r.bary[2] < fmin)
		fmin = r.bary[2];
	if (f2 < fmin)
		fmin = f2;
	return emin < 1e-4f || fmin < 1e-4f;
}

// intersection_resolver.cpp:841-870: a degenerate pair stays when one of its
// adjacent pairs is not degenerate.
__global__ void k_ct_keep(uint32_t P, const uint32_t* __restrict__ ids, const uint64_t* __restrict__ keys,
	const EfPairRec* __restrict__ pairs, const uint32_t* __restrict__ face_adj_edges,
	const uint32_t* __restrict__ edge_adj_faces, uint32_t* __restrict__ keep)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	bool k = !ef_degenerate(pairs[p]);
	if (!k)
		for_each_adjacent(p, P, ids, keys, face_adj_edges, edge_adj_faces,
			[&](uint32_t q)
			{
				k = !ef_degenerate(pairs[q]);
				return !k;
			});
	keep[p] = k ? 1u : 0u;
}

// Kept pairs, in order, into the second block of P slots of ids / pairs.
__global__ void k_ct_compact(uint32_t P, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
	uint32_t* __restrict__ ids, EfPairRec* __restrict__ pairs)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P || !keep[p])
		return;
	const uint32_t w = P + pos[p];
	ids[2 * w] = ids[2 * p];
	ids[2 * w + 1] = ids[2 * p + 1];
	pairs[w] = pairs[p];
}

__global__ void k_ct_keys(uint32_t P, const uint32_t* __restrict__ ids, uint64_t* __restrict__ keys)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p < P)
		keys[p] = ((uint64_t)ids[2 * p] << 32) | ids[2 * p + 1];
}

__global__ void k_ct_loop_count(uint32_t P, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ edges,
	const uint32_t* __restrict__ faces, const uint32_t* __restrict__ face_adj_edges,
	const uint32_t* __restrict__ edge_adj_faces, uint32_t* __restrict__ cnt)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	const uint32_t e = ids[2 * p], f = ids[2 * p + 1];
	uint32_t	   c = 0;
	for (int k = 0; k < 2; ++k)
	{
		const uint32_t T = edge_adj_faces[2 * e + k];
		uint32_t	   v;
		if (T != kNone && loop_vertex(e, f, T, edges, faces, face_adj_edges, &v) != kNone)
			++c;
	}
	cnt[p] = c;
}

// Neighbour counts of the P pairs (compute_adj_pairs plus their loop pairs).
__global__ void k_ct_adj_count(uint32_t P, const uint32_t* __restrict__ ids, const uint64_t* __restrict__ keys,
	const uint32_t* __restrict__ face_adj_edges, const uint32_t* __restrict__ edge_adj_faces,
	const uint32_t* __restrict__ loop_cnt, uint32_t* __restrict__ adj_n)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	uint32_t n = loop_cnt[p];
	for_each_adjacent(p, P, ids, keys, face_adj_edges, edge_adj_faces,
		[&](uint32_t)
		{
			++n;
			return true;
		});
	adj_n[p] = n;
}

// Loop pairs (intersection_resolver.cpp:1009-1018): pair P + r, r = loop_off[p] + j,
// edge bary (1, 0), face bary (1, 0, 0), all five indices the loop vertex; its
// only neighbour is its source pair, stored after the P pairs' neighbours.
__global__ void k_ct_loop_emit(uint32_t P, uint32_t* __restrict__ ids, EfPairRec* __restrict__ pairs,
	const uint32_t* __restrict__ edges, const uint32_t* __restrict__ faces,
	const uint32_t* __restrict__ face_adj_edges, const uint32_t* __restrict__ edge_adj_faces,
	const uint32_t* __restrict__ loop_off, const uint32_t* __restrict__ adj_total, uint32_t* __restrict__ adj,
	uint32_t* __restrict__ adj_off, uint32_t* __restrict__ adj_n)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	const uint32_t e = ids[2 * p], f = ids[2 * p + 1];
	const uint32_t base = *adj_total;
	uint32_t	   r = loop_off[p];
	for (int k = 0; k < 2; ++k)
	{
		const uint32_t T = edge_adj_faces[2 * e + k];
		uint32_t	   v = 0;
		const uint32_t le = T != kNone ? loop_vertex(e, f, T, edges, faces, face_adj_edges, &v) : kNone;
		if (le == kNone)
			continue;
		const uint32_t idx = P + r;
		ids[2 * idx] = le;
		ids[2 * idx + 1] = f;
		EfPairRec rec;
		for (int q = 0; q < 5; ++q)
			rec.idx[q] = v;
		rec.bary[0] = 1.0f;
		rec.bary[1] = 1.0f;
		rec.bary[2] = 0.0f;
		rec.grad[0] = rec.grad[1] = rec.grad[2] = 0.0f;
		rec.stiff = 0.0f;
		pairs[idx] = rec;
		adj[base + r] = p;
		adj_off[idx] = base + r;
		adj_n[idx] = 1;
		++r;
	}
}

__global__ void k_ct_adj_tail(uint32_t P, const uint32_t* __restrict__ loops, const uint32_t* __restrict__ adj_total,
	uint32_t* __restrict__ adj_off)
{
	if (threadIdx.x == 0)
		adj_off[P + *loops] = *adj_total + *loops;
}

__global__ void k_ct_adj_fill(uint32_t P, const uint32_t* __restrict__ ids, const uint64_t* __restrict__ keys,
	const uint32_t* __restrict__ face_adj_edges, const uint32_t* __restrict__ edge_adj_faces,
	const uint32_t* __restrict__ loop_cnt, const uint32_t* __restrict__ loop_off, const uint32_t* __restrict__ adj_off,
	uint32_t* __restrict__ adj)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	uint32_t* out = adj + adj_off[p];
	uint32_t  n = 0;
	for_each_adjacent(p, P, ids, keys, face_adj_edges, edge_adj_faces,
		[&](uint32_t q)
		{
			out[n++] = q;
			return true;
		});
	for (uint32_t j = 0; j < loop_cnt[p]; ++j)
		out[n++] = P + loop_off[p] + j;
}

__device__ __forceinline__ uint32_t flip_weight(const EfPairRec& a, const EfPairRec& b)
{
	return (a.idx[2] == b.idx[2] && a.idx[3] == b.idx[3] && a.idx[4] == b.idx[4]) ? 0u : 1u;
}

__global__ void k_ct_union(uint32_t N, const EfPairRec* __restrict__ pairs, const uint32_t* __restrict__ adj,
	const uint32_t* __restrict__ adj_off, uint32_t* parent)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= N)
		return;
	const EfPairRec rp = pairs[p];
	for (uint32_t s = adj_off[p]; s < adj_off[p + 1]; ++s)
	{
		const uint32_t q = adj[s];
		const uint32_t w = flip_weight(rp, pairs[q]);
		uf_union(parent, 2 * p, 2 * q + w);
		uf_union(parent, 2 * p + 1, 2 * q + (1u - w));
	}
}

__global__ void k_ct_contour_index(uint32_t N, const uint32_t* __restrict__ comp_min, const uint32_t* __restrict__ rank,
	uint32_t* __restrict__ out)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p < N)
		out[p] = rank[comp_min[p]];
}

// ---- contour classification ---------------------------------------------------
// classify_contour_type and the mesh_is_closed roles (intersection_resolver2.cpp:
// 311-375, :2811-2845) need, per contour, only counts: pairs on a boundary edge
// (at most one adjacent face), loop-vertex pairs (edge vertices equal), non-loop
// pairs with a single EF neighbour, per role the pairs and such degree-one pairs,
// and the meshes of the boundary pairs.  Counts are order-free, so one pass with
// atomics gives the reference's classification.
enum : uint32_t
{
	kCsBoundary = 0,
	kCsLoop = 1,
	kCsDeg1 = 2,
	kCsBoundaryMeshes = 3, // bit m: a boundary pair of mesh m
	kCsRolePairs = 4,	   // + role
	kCsRoleDeg1 = 6,	   // + role
	kCsWords = 8
};

__global__ void k_ct_classify_count(uint32_t N, const uint32_t* __restrict__ ids, const EfPairRec* __restrict__ pairs,
	const uint32_t* __restrict__ contour_index, const uint32_t* __restrict__ mesh_index,
	const uint32_t* __restrict__ adj_n, const uint32_t* __restrict__ edge_adj_faces, uint32_t* __restrict__ stats)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= N)
		return;
	uint32_t* const st = stats + (size_t)kCsWords * contour_index[p];
	const uint32_t	e = ids[2 * p], m = mesh_index[p] & 1u;
	const bool		boundary = edge_adj_faces[2 * e + 1] == kNone;
	const bool		loop = pairs[p].idx[0] == pairs[p].idx[1];
	const bool		deg1 = !loop && adj_n[p] == 1u;
	if (boundary)
	{
		atomicAdd(&st[kCsBoundary], 1u);
		atomicOr(&st[kCsBoundaryMeshes], 1u << m);
	}
	if (loop)
		atomicAdd(&st[kCsLoop], 1u);
	if (deg1)
	{
		atomicAdd(&st[kCsDeg1], 1u);
		atomicAdd(&st[kCsRoleDeg1 + m], 1u);
	}
	atomicAdd(&st[kCsRolePairs + m], 1u);
}

// ContourType (intersection_resolver2.h:27-37): 0 undefined, 1 closed, 2 eight,
// 3 LL, 4 BLI, 5 cross, 6 BIBI, 7 BBII.
__device__ __forceinline__ uint32_t contour_type(const uint32_t* st)
{
	const uint32_t nb = st[kCsBoundary], nl = st[kCsLoop];
	const bool	   closed = nb == 0u && st[kCsDeg1] == 0u;
	if (nb == 0u)
		return !closed ? 0u : nl == 0u ? 1u : nl == 1u ? 2u : nl == 2u ? 3u : 0u;
	if (nl == 0u)
		// two boundary pairs: on one mesh (BBII) or on both (BIBI)
		return nb == 2u ? (st[kCsBoundaryMeshes] == 3u ? 6u : 7u) : 0u;
	if (nl == 1u)
		return nb == 1u ? 4u : nb == 2u ? 5u : 0u;
	return 0u;
}

__global__ void k_ct_classify(uint32_t nc, const uint32_t* __restrict__ stats, uint32_t* __restrict__ type,
	uint32_t* __restrict__ mesh_closed)
{
	const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= nc)
		return;
	const uint32_t* st = stats + (size_t)kCsWords * c;
	type[c] = contour_type(st);
	// a role is closed when it owns a pair and none of its non-loop pairs has degree one
	for (uint32_t r = 0; r < 2; ++r)
		mesh_closed[2 * c + r] = (st[kCsRolePairs + r] > 0u && st[kCsRoleDeg1 + r] == 0u) ? 1u : 0u;
}

static int bits_for(uint32_t n)
{
	int b = 0;
	while (b < 32 && (n >> b) != 0u)
		++b;
	return b < 8 ? 8 : ((b + 7) / 8) * 8;
}

void ContourBuilder::init(Context* c, uint32_t nv_, uint32_t ne_, uint32_t nf_, const uint32_t* edges_,
	const uint32_t* faces_, const uint32_t* fae, const uint32_t* eaf)
{
	UT_REQUIRE(edges_ && faces_ && fae && eaf, "contour builder: null topology");
	ctx = c;
	nv = nv_;
	ne = ne_;
	nf = nf_;
	edges = edges_;
	faces = faces_;
	face_adj_edges = fae;
	edge_adj_faces = eaf;
	scalars.reserve(8);
	reserve(1024);
}

void ContourBuilder::reserve(uint32_t p)
{
	if (p <= cap)
		return;
	cap = p;
	const size_t total = 3 * (size_t)p; // up to two loop pairs per pair
	for (int k = 0; k < 2; ++k)
	{
		sort_k[k].reserve(p);
		sort_v[k].reserve(p);
	}
	sort_h.reserve(radix_sort_hist_size(p));
	scan_tmp.reserve(scan_tmp_size((uint32_t)total));
	keys.reserve(p);
	ids.reserve(2 * total);
	pairs.reserve(total);
	loop_cnt.reserve(p);
	loop_off.reserve(p);
	adj.reserve(8 * total);
	adj_off.reserve(total + 1);
	adj_n.reserve(total);
	uf.reserve(2 * total);
	comp_min.reserve(total);
	mesh_index.reserve(total);
	is_root.reserve(total);
	rank.reserve(total);
	contour_index.reserve(total);
	comp_inconsistent.reserve(total);
}

void ContourBuilder::build(uint32_t P, const EfPairRec* in_pairs, const uint32_t* in_ids, bool filter_degenerate)
{
	num_input = P;
	num_pairs = P;
	num_total = 0;
	num_adj = 0;
	num_contours = 0;
	if (P == 0)
		return;
	UT_REQUIRE(in_pairs && in_ids, "contour build: null input");
	reserve(P);
	hipStream_t s = ctx->stream;
	uint32_t	g = div_up(P, 256);
	// canonical order: stable sort by face, then by edge, then payload within equal keys
	RadixSortBuffers rs;
	for (int k = 0; k < 2; ++k)
	{
		rs.keys[k] = sort_k[k].get();
		rs.vals[k] = sort_v[k].get();
	}
	rs.hist = sort_h.get();
	k_ct_key_fid<<<g, 256, 0, s>>>(P, in_ids, rs.keys[0], rs.vals[0]);
	radix_sort_pairs(rs, P, bits_for(nf), s);
	if (rs.result != 0)
	{
		UT_HIP_CHECK(hipMemcpyAsync(rs.vals[0], rs.vals[1], 4 * (size_t)P, hipMemcpyDeviceToDevice, s));
	}
	k_ct_key_eid<<<g, 256, 0, s>>>(P, in_ids, rs.vals[0], rs.keys[0]);
	rs.result = 0;
	radix_sort_pairs(rs, P, bits_for(ne), s);
	const uint32_t* perm = rs.vals[rs.result];
	k_ct_gather<<<g, 256, 0, s>>>(P, perm, in_ids, in_pairs, ids.get(), pairs.get(), keys.get());
	k_ct_tie_order<<<g, 256, 0, s>>>(P, keys.get(), pairs.get());
	UT_LAUNCH_CHECK();
	uint32_t* d_loops = scalars.get() + 1;
	uint32_t* d_contours = scalars.get() + 2;
	uint32_t* d_adj = scalars.get() + 3;
	uint32_t* d_kept = scalars.get() + 4;
	if (filter_degenerate)
	{
		// the sort buffers are free once the pairs are gathered
		uint32_t* keep = sort_k[0].get();
		uint32_t* pos = sort_k[1].get();
		k_ct_keep<<<g, 256, 0, s>>>(P, ids.get(), keys.get(), pairs.get(), face_adj_edges, edge_adj_faces, keep);
		exclusive_scan(keep, pos, P, scan_tmp.get(), d_kept, s);
		UT_LAUNCH_CHECK();
		uint32_t K = 0;
		UT_HIP_CHECK(hipMemcpyAsync(&K, d_kept, 4, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		// no pair kept: all pairs stay (intersection_resolver.cpp:874-877)
		if (K > 0 && K < P)
		{
			k_ct_compact<<<g, 256, 0, s>>>(P, keep, pos, ids.get(), pairs.get());
			UT_HIP_CHECK(hipMemcpyAsync(ids.get(), ids.get() + 2 * (size_t)P, 8 * (size_t)K, hipMemcpyDeviceToDevice, s));
			UT_HIP_CHECK(hipMemcpyAsync(pairs.get(), pairs.get() + P, sizeof(EfPairRec) * K, hipMemcpyDeviceToDevice, s));
			P = K;
			g = div_up(P, 256);
			k_ct_keys<<<g, 256, 0, s>>>(P, ids.get(), keys.get());
			UT_LAUNCH_CHECK();
		}
		num_pairs = P;
	}
	// loop-vertex pairs and the packed adjacency
	k_ct_loop_count<<<g, 256, 0, s>>>(P, ids.get(), edges, faces, face_adj_edges, edge_adj_faces, loop_cnt.get());
	exclusive_scan(loop_cnt.get(), loop_off.get(), P, scan_tmp.get(), d_loops, s);
	k_ct_adj_count<<<g, 256, 0, s>>>(P, ids.get(), keys.get(), face_adj_edges, edge_adj_faces, loop_cnt.get(),
		adj_n.get());
	exclusive_scan(adj_n.get(), adj_off.get(), P, scan_tmp.get(), d_adj, s);
	k_ct_adj_tail<<<1, 64, 0, s>>>(P, d_loops, d_adj, adj_off.get());
	UT_LAUNCH_CHECK();
	uint32_t sc[5] = { 0, 0, 0, 0, 0 };
	UT_HIP_CHECK(hipMemcpyAsync(sc, scalars.get(), 20, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	const uint32_t L = sc[1], N = P + L;
	num_total = N;
	num_adj = sc[3] + L;
	adj.reserve(num_adj);
	k_ct_loop_emit<<<g, 256, 0, s>>>(P, ids.get(), pairs.get(), edges, faces, face_adj_edges, edge_adj_faces,
		loop_off.get(), d_adj, adj.get(), adj_off.get(), adj_n.get());
	k_ct_adj_fill<<<g, 256, 0, s>>>(P, ids.get(), keys.get(), face_adj_edges, edge_adj_faces, loop_cnt.get(),
		loop_off.get(), adj_off.get(), adj.get());
	// components + parity
	const uint32_t gn = div_up(N, 256), g2 = div_up(2 * N, 256);
	k_uf_init<<<g2, 256, 0, s>>>(uf.get(), 2 * N);
	k_fill_u32<<<gn, 256, 0, s>>>(comp_inconsistent.get(), N, 0u);
	k_ct_union<<<gn, 256, 0, s>>>(N, pairs.get(), adj.get(), adj_off.get(), uf.get());
	k_uf_compress<<<g2, 256, 0, s>>>(uf.get(), 2 * N);
	k_uf_labels<<<gn, 256, 0, s>>>(uf.get(), N, comp_min.get(), mesh_index.get(), is_root.get(),
		comp_inconsistent.get());
	exclusive_scan(is_root.get(), rank.get(), N, scan_tmp.get(), d_contours, s);
	k_ct_contour_index<<<gn, 256, 0, s>>>(N, comp_min.get(), rank.get(), contour_index.get());
	UT_LAUNCH_CHECK();
	uint32_t nc = 0;
	UT_HIP_CHECK(hipMemcpyAsync(&nc, d_contours, 4, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	num_contours = nc;
	// classification
	if (nc)
	{
		if (nc > cap_contours)
		{
			cap_contours = std::max(nc, 2 * cap_contours);
			contour_stats.reserve((size_t)kCsWords * cap_contours);
			contour_type.reserve(cap_contours);
			mesh_closed.reserve(2 * (size_t)cap_contours);
		}
		const uint32_t m = kCsWords * nc;
		k_fill_u32<<<div_up(m, 256), 256, 0, s>>>(contour_stats.get(), m, 0u);
		k_ct_classify_count<<<gn, 256, 0, s>>>(N, ids.get(), pairs.get(), contour_index.get(), mesh_index.get(),
			adj_n.get(), edge_adj_faces, contour_stats.get());
		k_ct_classify<<<div_up(nc, 256), 256, 0, s>>>(nc, contour_stats.get(), contour_type.get(), mesh_closed.get());
		UT_LAUNCH_CHECK();
	}
}

} // namespace ut
