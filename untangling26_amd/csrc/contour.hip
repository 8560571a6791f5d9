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

// Gather into canonical order (sorted by edge, then face; intersection_resolver.cpp:695-726).
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

__global__ void k_ct_loop_emit(uint32_t P, uint32_t* __restrict__ ids, EfPairRec* __restrict__ pairs,
	const uint32_t* __restrict__ edges, const uint32_t* __restrict__ faces,
	const uint32_t* __restrict__ face_adj_edges, const uint32_t* __restrict__ edge_adj_faces,
	const uint32_t* __restrict__ off, uint32_t* __restrict__ adj, uint32_t* __restrict__ adj_n)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	const uint32_t e = ids[2 * p], f = ids[2 * p + 1];
	uint32_t	   idx = P + off[p];
	for (int k = 0; k < 2; ++k)
	{
		const uint32_t T = edge_adj_faces[2 * e + k];
		uint32_t	   v = 0;
		const uint32_t le = T != kNone ? loop_vertex(e, f, T, edges, faces, face_adj_edges, &v) : kNone;
		if (le == kNone)
			continue;
		ids[2 * idx] = le;
		ids[2 * idx + 1] = f;
		EfPairRec r;
		for (int q = 0; q < 5; ++q)
			r.idx[q] = v;
		r.bary[0] = 1.0f; // edge bary (1, 0), face bary (1, 0, 0)
		r.bary[1] = 1.0f;
		r.bary[2] = 0.0f;
		r.grad[0] = r.grad[1] = r.grad[2] = 0.0f;
		r.stiff = 0.0f;
		pairs[idx] = r;
		adj[kMaxEfAdj * idx] = p;
		adj_n[idx] = 1;
		++idx;
	}
}

__device__ __forceinline__ int find_pair(const uint64_t* __restrict__ keys, uint32_t P, uint32_t e, uint32_t f)
{
	const uint64_t k = ((uint64_t)e << 32) | f;
	uint32_t	   lo = 0, hi = P;
	while (lo < hi)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (keys[mid] < k)
			lo = mid + 1;
		else
			hi = mid;
	}
	return (lo < P && keys[lo] == k) ? (int)lo : -1;
}

// compute_adj_pairs (intersection_resolver.cpp:503-558) plus the loop pairs
// appended to their source pair (:1009-1018), in the reference's order.
__global__ void k_ct_adjacency(uint32_t P, const uint32_t* __restrict__ ids, const uint64_t* __restrict__ keys,
	const uint32_t* __restrict__ face_adj_edges, const uint32_t* __restrict__ edge_adj_faces,
	const uint32_t* __restrict__ loop_cnt, const uint32_t* __restrict__ loop_off, uint32_t* __restrict__ adj,
	uint32_t* __restrict__ adj_n, uint32_t* __restrict__ status)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= P)
		return;
	const uint32_t e = ids[2 * p], f = ids[2 * p + 1];
	const uint32_t T[2] = { edge_adj_faces[2 * e], edge_adj_faces[2 * e + 1] };
	uint32_t	   n = 0;
	auto		   push = [&](uint32_t q)
	{
		if (n < kMaxEfAdj)
			adj[kMaxEfAdj * p + n] = q;
		++n;
	};
	// via the other edges of the faces around e, same intersected face
	for (int k = 0; k < 2; ++k)
	{
		if (T[k] == kNone)
			continue;
		const uint32_t a = face_adj_edges[3 * T[k]], b = face_adj_edges[3 * T[k] + 1], c = face_adj_edges[3 * T[k] + 2];
		const uint32_t o0 = (a == e) ? b : a;
		const uint32_t o1 = (a == e || b == e) ? c : b;
		const int	   q0 = find_pair(keys, P, o0, f);
		if (q0 >= 0 && (uint32_t)q0 != p)
			push((uint32_t)q0);
		const int q1 = find_pair(keys, P, o1, f);
		if (q1 >= 0 && (uint32_t)q1 != p)
			push((uint32_t)q1);
	}
	// via an edge of the intersected face cutting a face around e
	for (int ii = 0; ii < 3; ++ii)
	{
		const uint32_t fe = face_adj_edges[3 * f + ii];
		if (fe == e)
			continue;
		for (int k = 0; k < 2; ++k)
		{
			if (T[k] == kNone)
				continue;
			const int q = find_pair(keys, P, fe, T[k]);
			if (q >= 0 && (uint32_t)q != p)
				push((uint32_t)q);
		}
	}
	for (uint32_t j = 0; j < loop_cnt[p]; ++j)
		push(P + loop_off[p] + j);
	if (n > kMaxEfAdj)
	{
		atomicAdd(&status[0], 1u);
		n = kMaxEfAdj;
	}
	adj_n[p] = n;
}

__device__ __forceinline__ uint32_t flip_weight(const EfPairRec& a, const EfPairRec& b)
{
	return (a.idx[2] == b.idx[2] && a.idx[3] == b.idx[3] && a.idx[4] == b.idx[4]) ? 0u : 1u;
}

__global__ void k_ct_union(uint32_t N, const EfPairRec* __restrict__ pairs, const uint32_t* __restrict__ adj,
	const uint32_t* __restrict__ adj_n, uint32_t* parent)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= N)
		return;
	const EfPairRec rp = pairs[p];
	for (uint32_t s = 0; s < adj_n[p]; ++s)
	{
		const uint32_t q = adj[kMaxEfAdj * p + s];
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
	const size_t total = 3 * (size_t)p; // up to two loop pairs per DCD pair
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
	adj.reserve(kMaxEfAdj * total);
	adj_n.reserve(total);
	uf.reserve(2 * total);
	comp_min.reserve(total);
	mesh_index.reserve(total);
	is_root.reserve(total);
	rank.reserve(total);
	contour_index.reserve(total);
	comp_inconsistent.reserve(total);
}

void ContourBuilder::build(uint32_t P, const EfPairRec* in_pairs, const uint32_t* in_ids)
{
	num_pairs = P;
	num_total = 0;
	num_contours = 0;
	if (P == 0)
		return;
	UT_REQUIRE(in_pairs && in_ids, "contour build: null input");
	reserve(P);
	hipStream_t s = ctx->stream;
	const uint32_t g = div_up(P, 256);
	// canonical order: stable sort by face, then by edge
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
	// loop-vertex pairs
	uint32_t* d_status = scalars.get();
	uint32_t* d_loops = scalars.get() + 1;
	uint32_t* d_contours = scalars.get() + 2;
	k_fill_u32<<<1, 64, 0, s>>>(d_status, 1, 0u);
	k_ct_loop_count<<<g, 256, 0, s>>>(P, ids.get(), edges, faces, face_adj_edges, edge_adj_faces, loop_cnt.get());
	exclusive_scan(loop_cnt.get(), loop_off.get(), P, scan_tmp.get(), d_loops, s);
	k_ct_loop_emit<<<g, 256, 0, s>>>(P, ids.get(), pairs.get(), edges, faces, face_adj_edges, edge_adj_faces,
		loop_off.get(), adj.get(), adj_n.get());
	k_ct_adjacency<<<g, 256, 0, s>>>(P, ids.get(), keys.get(), face_adj_edges, edge_adj_faces, loop_cnt.get(),
		loop_off.get(), adj.get(), adj_n.get(), d_status);
	UT_LAUNCH_CHECK();
	uint32_t L = 0;
	UT_HIP_CHECK(hipMemcpyAsync(&L, d_loops, 4, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	const uint32_t N = P + L;
	num_total = N;
	const uint32_t gn = div_up(N, 256), g2 = div_up(2 * N, 256);
	// components + parity
	k_uf_init<<<g2, 256, 0, s>>>(uf.get(), 2 * N);
	k_fill_u32<<<gn, 256, 0, s>>>(comp_inconsistent.get(), N, 0u);
	k_ct_union<<<gn, 256, 0, s>>>(N, pairs.get(), adj.get(), adj_n.get(), uf.get());
	k_uf_compress<<<g2, 256, 0, s>>>(uf.get(), 2 * N);
	k_uf_labels<<<gn, 256, 0, s>>>(uf.get(), N, comp_min.get(), mesh_index.get(), is_root.get(),
		comp_inconsistent.get());
	exclusive_scan(is_root.get(), rank.get(), N, scan_tmp.get(), d_contours, s);
	k_ct_contour_index<<<gn, 256, 0, s>>>(N, comp_min.get(), rank.get(), contour_index.get());
	UT_LAUNCH_CHECK();
	uint32_t sc[3] = { 0, 0, 0 };
	UT_HIP_CHECK(hipMemcpyAsync(sc, scalars.get(), 12, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	num_contours = sc[2];
	if (sc[0] != 0)
		throw Error(kInternal, "EF pair adjacency exceeded " + std::to_string(kMaxEfAdj) + " entries");
}

} // namespace ut
