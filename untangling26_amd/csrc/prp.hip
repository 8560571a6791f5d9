// PRP untangling on gfx950 (see prp.hpp for the split between host and device).
//
// Reference semantics: intersection_resolver2.cpp (CPU PRP, the default
// resolver) with the ray tests of cached_bvh.cpp:131-555 and the hit
// adjacency of intersection_resolver_helper.cpp:771-902.  Kernels are FP32 with
// no contraction, so every ray test rounds like the reference's scalar code.
#include "prp.hpp"

#include "lbvh_traverse.hpp"
#include "radix_sort.hpp"
#include "scan.hpp"
#include "seg_tri.hpp"
#include "ut_common.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace ut
{

namespace
{

constexpr uint32_t kHopInf = 0xFFFFFFFFu;
constexpr uint32_t kOwnUnknown = 0xFFu;
constexpr int	   kBfsThreads = 1024;
// Persistent ray grid: 6 workgroups per CU on 256 CUs (the ray kernels use ~80
// VGPRs, so 6 waves per SIMD), each lane with its own spill stack.  A 16-entry
// LDS stack keeps 6 workgroups within the 160 KB of LDS per CU.
constexpr int	   kRayBlocks = 1536;
// the EE sweep runs at 5 waves per SIMD (its leaf tests keep the request's constants in
// registers): one resident workgroup of 4 waves per 5 / 4 SIMD slots of each of 256 CUs
constexpr int	   kEeWavesPerSimd = 5;
constexpr int	   kEeBlocks = 256 * kEeWavesPerSimd;
static_assert(kEeBlocks <= kRayBlocks && kEeBlocks % kNumXcd == 0, "the EE grid uses the ray spill stacks");
constexpr int	   kRayStackLds = 16;
static_assert(kRayBlocks % kNumXcd == 0, "the ray grid is split evenly over the XCDs");
constexpr uint32_t kMaxDirs = 64;
constexpr uint64_t kHashEmpty = ~0ull;

// ------------------------------------------------------------------ normals (:2605-2644)
__device__ __forceinline__ f3 safe_normalize(f3 v)
{
	const float l2 = dot3(v, v);
	return l2 > 1e-20f ? v * (1.0f / sqrtf(l2)) : mk3(0.0f, 0.0f, 0.0f);
}

__global__ void k_prp_face_normals(uint32_t nf, const float* __restrict__ x, const uint32_t* __restrict__ faces,
	float* __restrict__ fn)
{
	const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
	if (f >= nf)
		return;
	const f3 a = ld3(x, faces[3 * f]), b = ld3(x, faces[3 * f + 1]), c = ld3(x, faces[3 * f + 2]);
	st3(fn, f, safe_normalize(cross3(b - a, c - a)));
}

// edge normals (area-weighted faces of the edge), then vertex normals (faces of the vertex, ascending)
__global__ void k_prp_edge_vert_normals(uint32_t ne, uint32_t nv, const uint32_t* __restrict__ eaf,
	const uint32_t* __restrict__ vf_off, const uint32_t* __restrict__ vf, const float* __restrict__ fa,
	const float* __restrict__ fn, float* __restrict__ en, float* __restrict__ vn)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ne)
	{
		f3 s = mk3(0.0f, 0.0f, 0.0f);
		for (int k = 0; k < 2; ++k)
		{
			const uint32_t f = eaf[2 * i + k];
			if (f != kNone)
				s = s + ld3(fn, f) * fa[f];
		}
		st3(en, i, safe_normalize(s));
	}
	else if (i < ne + nv)
	{
		const uint32_t v = i - ne;
		f3			   s = mk3(0.0f, 0.0f, 0.0f);
		for (uint32_t k = vf_off[v]; k < vf_off[v + 1]; ++k)
			s = s + ld3(fn, vf[k]) * fa[vf[k]];
		st3(vn, v, safe_normalize(s));
	}
}

// ------------------------------------------------------------------ distances (:618-720)
// Neighbour blocks: for every vertex its first eight neighbours with their rest
// edge lengths, (neighbour, length bits) pairs padded with (kNone, 0), 64 bytes
// per vertex so one vertex's block is a few 16-byte loads.  Lengths use the
// expression of the relaxation itself, so the sums round identically.
constexpr int kAdjWays = 8;

__device__ __forceinline__ float rest_len(const float* __restrict__ rest_x, f3 xu, uint32_t w)
{
	const f3 d = xu - ld3(rest_x, w);
	return sqrtf(dot3(d, d));
}

__global__ void k_prp_adj_blocks(uint32_t V, const uint32_t* __restrict__ vv_off, const uint32_t* __restrict__ vv,
	const float* __restrict__ rest_x, uint2* __restrict__ adj)
{
	const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
	if (u >= V)
		return;
	const uint32_t k0 = vv_off[u], deg = vv_off[u + 1] - k0;
	const f3	   xu = ld3(rest_x, u);
	for (uint32_t q = 0; q < kAdjWays; ++q)
	{
		uint2 e = make_uint2(kNone, 0u);
		if (q < deg)
		{
			const uint32_t w = vv[k0 + q];
			e = make_uint2(w, __float_as_uint(rest_len(rest_x, xu, w)));
		}
		adj[(size_t)kAdjWays * u + q] = e;
	}
}

// One workgroup per listed (contour, role) slot: the BFS hop distance and the
// Bellman-Ford geodesic distance on the rest edge lengths, both as
// label-correcting relaxations over one work list.  Each fixpoint is order
// independent: the hop count is the minimum over paths of the edge count, and
// the geodesic the minimum over paths of the path's left-to-right rounded
// FP32 length (what Dijkstra's FP32 relaxations give), since rounding is
// monotone.
//
// The search is truncated and resumable: a vertex is expanded only while its
// hop label is <= the slot's hop threshold k or its geodesic label <= the
// slot's radius r.  Every vertex whose true hop is <= k + 1 then has its exact
// hop (the BFS path's vertices all expand), every vertex whose true geodesic
// is <= r its exact geodesic (a shortest path's prefixes are no longer), and
// every other label is an upper bound above those limits.  Raising the
// thresholds resumes the search from the vertices that now qualify and had
// not (the labels stay in the slot's scratch).  Mode A raises k (from a
// request) and reports the largest finite hop label, which is the diameter
// exactly when it is <= k (the BFS died out), and the largest geodesic label
// on the ring hop <= k, an upper bound of the ring's true geodesics.  Mode B
// raises r to the larger of the two roles' ring bounds, after which every
// geodesic that the ownership test compares is exact or provably above the
// vertex's own distance.  Infinite thresholds give the full fields.
//
// A vertex is listed for the next round when either label improves, once per
// round: the "listed" marks are a bitmap in LDS per list where the mesh allows
// (two alternating lists; a list's bits are cleared as its entries are taken,
// so it is clean when it is filled again), else generation-stamped words in
// global memory.  The first entries of each list sit in LDS, the rest in
// global memory.  The (hop, geodesic) labels of a vertex are one 8-byte word
// of the slot's scratch, unpacked into the hop and geodesic rows after each
// launch.
constexpr uint32_t kDistListLds = 2048;	  // list entries per round held in LDS
constexpr uint32_t kDistLdsWords = 16384; // 64 KB of LDS per workgroup: two lists + two bitmaps
constexpr uint32_t kDistHopAll = 0xFFFFFFFEu; // hop threshold that expands every reached vertex

uint32_t dist_bitmap_words(uint32_t V)
{
	const uint32_t w = (V + 31) / 32;
	return 2 * kDistListLds + 2 * w <= kDistLdsWords ? w : 0u;
}

struct DistArgs
{
	uint32_t		V, c0;
	const uint32_t* req;   // mode A: (slot, hop threshold, first launch) per workgroup; mode B: (slot, -, -)
	uint32_t		mode_b;
	const uint32_t *b_off, *b_verts, *vv_off, *vv;
	const float*	rest_x;
	const uint2*	adj;
	uint32_t*		hop_all; // [slot][V] rows
	float*			geo_all;
	uint32_t*		queue_all; // [slot][4V]: list overflow [2][V], then the label words [V] (8 bytes each)
	uint32_t*		inq_all;   // [slot][V] listed stamps (meshes beyond the LDS bitmap)
	uint2*			state;	   // [slot] (hop threshold, radius bits) reached so far
	uint2*			out;	   // [slot] (largest finite hop, ring geodesic bound bits), written by mode A
	uint32_t		bitmap_words;
};

__global__ __launch_bounds__(kBfsThreads) void k_prp_distances(DistArgs a)
{
	extern __shared__ uint32_t dsm[];
	const uint32_t V = a.V;
	const uint32_t s = a.req[3 * blockIdx.x]; // slot: 2 * local contour + role
	const bool	   first = !a.mode_b && a.req[3 * blockIdx.x + 2] != 0;
	const uint32_t bs = 2 * a.c0 + s;		  // global (contour, role) slot
	uint32_t*	   Q = a.queue_all + (size_t)s * 4 * V;
	uint2*		   HG = reinterpret_cast<uint2*>(Q + 2 * (size_t)V);
	uint32_t*	   hg = Q + 2 * (size_t)V; // the same words as 32-bit halves: hop 2v, geodesic 2v + 1
	uint32_t*	   I = a.inq_all + (size_t)s * V;
	const bool	   bitmap = a.bitmap_words > 0;
	// thresholds before and after this launch (a first launch has none before)
	const uint2	   old = first ? make_uint2(0u, 0u) : a.state[s];
	uint32_t	   kn = old.x, rn = old.y;
	if (!a.mode_b)
		kn = max(kn, a.req[3 * blockIdx.x + 1]);
	else
		rn = max(rn, max(a.out[s].y, a.out[s ^ 1u].y)); // non-negative float bits order as the values
	auto qual_old = [&](uint32_t h, uint32_t g) { return !first && (h <= old.x || g <= old.y); };
	auto qual_new = [&](uint32_t h, uint32_t g) { return h <= kn || g <= rn; };
	__shared__ uint32_t nnext, n0, hmax, gmax;
	if (threadIdx.x == 0)
		nnext = 0, n0 = 0, hmax = 0, gmax = 0;
	for (uint32_t k = threadIdx.x; k < 2 * a.bitmap_words; k += blockDim.x)
		dsm[2 * kDistListLds + k] = 0u;
	if (!bitmap)
		for (uint32_t v = threadIdx.x; v < V; v += blockDim.x)
			I[v] = 0u;
	// entry i of list l
	auto entry = [&](uint32_t l, uint32_t i) -> uint32_t&
	{ return i < kDistListLds ? dsm[l * kDistListLds + i] : Q[(size_t)l * V + (i - kDistListLds)]; };
	if (first)
	{
		for (uint32_t v = threadIdx.x; v < V; v += blockDim.x)
			HG[v] = make_uint2(kHopInf, __float_as_uint(FLT_MAX));
		__syncthreads();
		const uint32_t b0 = a.b_off[bs], nb = a.b_off[bs + 1] - b0;
		for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
		{
			const uint32_t v = a.b_verts[b0 + i];
			HG[v] = make_uint2(0u, 0u);
			entry(0, i) = v;
		}
		if (threadIdx.x == 0)
			n0 = nb;
	}
	else
	{
		__syncthreads();
		// resume from the reached vertices that qualify now and did not before
		for (uint32_t v = threadIdx.x; v < V; v += blockDim.x)
		{
			const uint2 e = HG[v];
			if (e.x != kHopInf && !qual_old(e.x, e.y) && qual_new(e.x, e.y))
				entry(0, atomicAdd(&n0, 1u)) = v;
		}
	}
	__syncthreads();
	uint32_t ncur = n0, gen = 1, cl = 0; // entries, round stamp, list being taken
	// relaxation of the edge u -> w (w != kNone) with rest length bits lb; cw is a read of w's word
	auto relax_edge = [&](uint32_t hu, float gu, uint32_t w, uint32_t lb, uint2 cw)
	{
		const uint32_t nh = hu == kHopInf ? kHopInf : hu + 1;
		const uint32_t ng = __float_as_uint(gu + __uint_as_float(lb));
		// values only decrease, so a read that is not above the candidate (stale or
		// not) proves the atomic would change nothing
		bool better = false;
		if (nh < cw.x)
			better = nh < atomicMin(&hg[2 * w], nh);
		if (ng < cw.y)
			better = (ng < atomicMin(&hg[2 * w + 1], ng)) || better;
		if (!better)
			return;
		bool add;
		if (bitmap)
		{
			const uint32_t bit = 1u << (w & 31);
			add = !(atomicOr(&dsm[2 * kDistListLds + (cl ^ 1) * a.bitmap_words + (w >> 5)], bit) & bit);
		}
		else
			add = atomicExch(&I[w], gen) != gen;
		if (add)
			entry(cl ^ 1, atomicAdd(&nnext, 1u)) = w;
	};
	while (ncur > 0)
	{
		if (threadIdx.x == 0)
			nnext = 0;
		__syncthreads();
		for (uint32_t i = threadIdx.x; i < ncur; i += blockDim.x)
		{
			const uint32_t u = entry(cl, i);
			if (bitmap) // u leaves the list: its mark is cleared for the list's next filling
				atomicAnd(&dsm[2 * kDistListLds + cl * a.bitmap_words + (u >> 5)], ~(1u << (u & 31)));
			const uint64_t wu = __hip_atomic_load(
				reinterpret_cast<const uint64_t*>(HG) + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
			const uint32_t hu = (uint32_t)wu;
			const float	   gu = __uint_as_float((uint32_t)(wu >> 32));
			if (!qual_new(hu, (uint32_t)(wu >> 32)))
				continue; // beyond both thresholds: waits for a later launch
			const uint4* blk = reinterpret_cast<const uint4*>(a.adj + (size_t)kAdjWays * u);
			uint32_t	 w[kAdjWays], lb[kAdjWays];
#pragma unroll
			for (int q = 0; q < kAdjWays / 2; ++q)
			{
				const uint4 b = blk[q];
				w[2 * q] = b.x, lb[2 * q] = b.y, w[2 * q + 1] = b.z, lb[2 * q + 1] = b.w;
			}
			uint2 cw[kAdjWays];
#pragma unroll
			for (int q = 0; q < kAdjWays; ++q)
				cw[q] = w[q] != kNone ? HG[w[q]] : make_uint2(0u, 0u);
#pragma unroll
			for (int q = 0; q < kAdjWays; ++q)
				if (w[q] != kNone)
					relax_edge(hu, gu, w[q], lb[q], cw[q]);
			// neighbours beyond the block (degree above eight)
			const uint32_t k1 = a.vv_off[u + 1];
			if (k1 - a.vv_off[u] > kAdjWays)
			{
				const f3 xu = ld3(a.rest_x, u);
				for (uint32_t k = a.vv_off[u] + kAdjWays; k < k1; ++k)
				{
					const uint32_t x = a.vv[k];
					relax_edge(hu, gu, x, __float_as_uint(rest_len(a.rest_x, xu, x)), HG[x]);
				}
			}
		}
		__syncthreads();
		ncur = nnext;
		cl ^= 1;
		++gen;
		__syncthreads();
	}
	// unpack into the hop and geodesic rows: the largest finite hop label and the
	// largest geodesic label on the ring hop <= k
	uint32_t* H = a.hop_all + (size_t)s * V;
	float*	  G = a.geo_all + (size_t)s * V;
	uint32_t  mh = 0, mg = 0;
	for (uint32_t v = threadIdx.x; v < V; v += blockDim.x)
	{
		const uint64_t e = __hip_atomic_load(
			reinterpret_cast<const uint64_t*>(HG) + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		const uint32_t h = (uint32_t)e, g = (uint32_t)(e >> 32);
		H[v] = h, G[v] = __uint_as_float(g);
		if (h != kHopInf)
			mh = max(mh, h);
		if (h <= kn)
			mg = max(mg, g);
	}
	atomicMax(&hmax, mh);
	atomicMax(&gmax, mg);
	__syncthreads();
	if (threadIdx.x == 0)
	{
		a.state[s] = make_uint2(kn, rn);
		if (!a.mode_b)
			a.out[s] = make_uint2(hmax, gmax);
	}
}

// ------------------------------------------------------------------ ownership + candidates (:1571-1804)
__device__ __forceinline__ float normalized_phi(float own, float opp)
{
	const float unreachable = FLT_MAX * 0.5f;
	const bool	g0 = isfinite(own) && own < unreachable, g1 = isfinite(opp) && opp < unreachable;
	if (!g0 && !g1)
		return 0.0f;
	if (!g0)
		return 1.0f;
	if (!g1)
		return -1.0f;
	return (own - opp) / (own + opp + 1e-6f);
}

__device__ __forceinline__ bool owns1(float own, float opp, float tau)
{
	const float unreachable = FLT_MAX * 0.5f;
	if (!isfinite(own) || own >= unreachable)
		return false;
	if (!isfinite(opp) || opp >= unreachable)
		return true;
	return normalized_phi(own, opp) <= tau;
}

__device__ __forceinline__ float min_dist_to(f3 p, const float* __restrict__ x, const uint32_t* __restrict__ b,
	uint32_t n)
{
	float best = FLT_MAX;
	for (uint32_t i = 0; i < n; ++i)
	{
		const f3	d = p - ld3(x, b[i]);
		const float l = sqrtf(dot3(d, d));
		best = l < best ? l : best;
	}
	return best;
}

// Ownership of every vertex within hop k of its role (computed once per vertex;
// k only grows), then the candidate vertex and edge lists of each (contour, role).
__global__ void k_prp_own(uint32_t V, uint32_t c0, uint32_t nslots2, const uint32_t* __restrict__ kring,
	const uint32_t* __restrict__ hop_all, const float* __restrict__ geo_all, const float* __restrict__ x,
	const uint32_t* __restrict__ b_off, const uint32_t* __restrict__ b_verts, float tau, uint8_t* __restrict__ own_all)
{
	const uint32_t s = blockIdx.y;
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nslots2 || v >= V)
		return;
	const size_t o = (size_t)s * V + v;
	if (hop_all[o] > kring[s] || own_all[o] != kOwnUnknown)
		return;
	const uint32_t so = s ^ 1u;
	const uint32_t bs = 2 * c0 + s, bo = 2 * c0 + so;
	const f3	   p = ld3(x, v);
	const float	   cur_own = min_dist_to(p, x, b_verts + b_off[bs], b_off[bs + 1] - b_off[bs]);
	const float	   cur_opp = min_dist_to(p, x, b_verts + b_off[bo], b_off[bo + 1] - b_off[bo]);
	const float	   rest_own = geo_all[o], rest_opp = geo_all[(size_t)so * V + v];
	own_all[o] = (uint8_t)(owns1(rest_own, rest_opp, tau) && owns1(cur_own, cur_opp, tau));
}

__device__ __forceinline__ bool cand_vert(const uint32_t* H, const uint8_t* O, uint32_t k, uint32_t v)
{
	return H[v] <= k && O[v] == 1u;
}
__device__ __forceinline__ bool cand_edge(const uint32_t* H, const uint8_t* O, uint32_t k, uint32_t a, uint32_t b)
{
	return H[a] <= k && H[b] <= k && (O[a] == 1u || O[b] == 1u);
}
__device__ __forceinline__ bool cand_face(const uint32_t* H, const uint8_t* O, uint32_t k, uint32_t a, uint32_t b,
	uint32_t c)
{
	return H[a] <= k && H[b] <= k && H[c] <= k && (O[a] == 1u || O[b] == 1u || O[c] == 1u);
}

// cand lists: per slot s, vertices at [s * (V + E)], edges at [s * (V + E) + V]
// Also the candidate edges as a bitmap, ecand[s * ewords + e / 64] bit e % 64
// (grid rows of 64 lanes are aligned to 64 edges, one word per wave).
__global__ void k_prp_cand_lists(uint32_t V, uint32_t E, uint32_t nslots2, const uint32_t* __restrict__ kring,
	const uint32_t* __restrict__ hop_all, const uint8_t* __restrict__ own_all, const uint32_t* __restrict__ edges,
	uint32_t* __restrict__ cand, uint32_t* __restrict__ cand_cnt, uint64_t* __restrict__ ecand, uint32_t ewords)
{
	const uint32_t s = blockIdx.y;
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nslots2)
		return;
	const uint32_t* H = hop_all + (size_t)s * V;
	const uint8_t*	O = own_all + (size_t)s * V;
	const uint32_t	k = kring[s];
	uint32_t*		C = cand + (size_t)s * (V + E);
	const bool		vin = i < V && cand_vert(H, O, k, i);
	const uint32_t	pv = wave_append(&cand_cnt[2 * s], vin);
	if (vin)
		C[pv] = i;
	const bool	   ein = i < E && cand_edge(H, O, k, edges[2 * i], edges[2 * i + 1]);
	const uint32_t pe = wave_append(&cand_cnt[2 * s + 1], ein);
	if (ein)
		C[V + pe] = i;
	const uint64_t em = __ballot(ein);
	if (__lane_id() == 0 && (i >> 6) < ewords)
		ecand[(size_t)s * ewords + (i >> 6)] = em;
}

// Per slot: projections of the candidate target edges (role 1) on the slot
// direction (intersection_resolver2.h:572-595), and the boxes of the candidate
// edges of both roles: tbox[12 j .. +5] role 1, [12 j + 6 .. +11] role 0.  Every
// candidate face has two candidate edges (at its owned vertex) covering its
// corners, so these boxes also hold the candidate faces of the role.
__global__ void k_prp_slot_prep(uint32_t V, uint32_t E, uint32_t nslots, PrpSlot* __restrict__ slots,
	const uint32_t* __restrict__ cand, const uint32_t* __restrict__ cand_cnt, const float* __restrict__ x,
	const uint32_t* __restrict__ edges, float* __restrict__ tbox)
{
	const uint32_t j = blockIdx.x;
	if (j >= nslots)
		return;
	PrpSlot		   sl = slots[j];
	const f3	   dir = mk3(sl.dir[0], sl.dir[1], sl.dir[2]);
	// [0] projection min, [1] projection max, [2..4] box min, [5..7] box max of the
	// role-1 endpoints; [8..10] box min, [11..13] box max of the role-0 endpoints
	float r[14] = { FLT_MAX, -FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, FLT_MAX, FLT_MAX,
		FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX };
	for (uint32_t role = 0; role < 2; ++role)
	{
		const uint32_t	s_ = 2 * sl.contour + role;
		const uint32_t	n = cand_cnt[2 * s_ + 1];
		const uint32_t* C = cand + (size_t)s_ * (V + E) + V;
		float*			b = role ? r + 2 : r + 8;
		for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
		{
			const uint32_t e = C[i];
			const f3	   q0 = ld3(x, edges[2 * e]), q1 = ld3(x, edges[2 * e + 1]);
			if (role)
			{
				const float p0 = dot3(q0, dir), p1 = dot3(q1, dir);
				r[0] = fminf(r[0], fminf(p0, p1));
				r[1] = fmaxf(r[1], fmaxf(p0, p1));
			}
			b[0] = fminf(b[0], fminf(q0.x, q1.x)), b[1] = fminf(b[1], fminf(q0.y, q1.y));
			b[2] = fminf(b[2], fminf(q0.z, q1.z));
			b[3] = fmaxf(b[3], fmaxf(q0.x, q1.x)), b[4] = fmaxf(b[4], fmaxf(q0.y, q1.y));
			b[5] = fmaxf(b[5], fmaxf(q0.z, q1.z));
		}
	}
	auto is_min = [](int k) { return k == 0 || (k >= 2 && k <= 4) || (k >= 8 && k <= 10); };
	__shared__ float sr[14][16];
	const uint32_t	 w = threadIdx.x / kWave;
	for (int k = 0; k < 14; ++k)
	{
		const float v = is_min(k) ? wave_min(r[k]) : wave_max(r[k]);
		if (__lane_id() == 0)
			sr[k][w] = v;
	}
	__syncthreads();
	if (threadIdx.x < 14)
	{
		const int k = threadIdx.x;
		float	  a = sr[k][0];
		for (uint32_t q = 1; q < blockDim.x / kWave; ++q)
			a = is_min(k) ? fminf(a, sr[k][q]) : fmaxf(a, sr[k][q]);
		const uint32_t n1 = cand_cnt[2 * (2 * sl.contour + 1) + 1];
		if (k == 0)
			slots[j].ee_pmin = n1 ? a : 0.0f;
		else if (k == 1)
			slots[j].ee_pmax = n1 ? a : 0.0f;
		else
			tbox[12 * (size_t)j + (k - 2)] = a; // empty candidate set: an inverted box that meets no node
	}
}

// request counts per slot: [VF, EE, FV] = (|cand verts role 0|, |cand edges role 0|, |cand verts role 1|),
// a kind is skipped when its target set is empty (intersection_resolver2.cpp:3040-3056)
__global__ void k_prp_request_offsets(uint32_t nslots, const PrpSlot* __restrict__ slots,
	const uint32_t* __restrict__ cand_cnt, const uint32_t* __restrict__ target_cnt, uint32_t* __restrict__ off)
{
	// single workgroup of 1024 threads, nslots <= 4096
	__shared__ uint32_t part[3][1024];
	const uint32_t		per = (nslots + blockDim.x - 1) / blockDim.x;
	uint32_t			loc[3] = { 0, 0, 0 };
	for (uint32_t q = 0; q < per; ++q)
	{
		const uint32_t j = threadIdx.x * per + q;
		if (j >= nslots)
			break;
		const uint32_t c = slots[j].contour;
		const uint32_t s0 = 2 * c, s1 = 2 * c + 1;
		loc[0] += target_cnt[s1] ? cand_cnt[2 * s0] : 0u;			 // VF needs target faces of role 1
		loc[1] += cand_cnt[2 * s1 + 1] ? cand_cnt[2 * s0 + 1] : 0u; // EE needs target edges of role 1
		loc[2] += target_cnt[s0] ? cand_cnt[2 * s1] : 0u;			 // FV needs source faces of role 0
	}
	for (int t = 0; t < 3; ++t)
		part[t][threadIdx.x] = loc[t];
	__syncthreads();
	if (threadIdx.x < 3)
	{
		uint32_t acc = 0;
		for (uint32_t i = 0; i < blockDim.x; ++i)
		{
			const uint32_t v = part[threadIdx.x][i];
			part[threadIdx.x][i] = acc;
			acc += v;
		}
		off[threadIdx.x * (nslots + 1) + nslots] = acc;
	}
	__syncthreads();
	uint32_t run[3] = { part[0][threadIdx.x], part[1][threadIdx.x], part[2][threadIdx.x] };
	for (uint32_t q = 0; q < per; ++q)
	{
		const uint32_t j = threadIdx.x * per + q;
		if (j >= nslots)
			break;
		const uint32_t c = slots[j].contour;
		const uint32_t s0 = 2 * c, s1 = 2 * c + 1;
		const uint32_t n[3] = { target_cnt[s1] ? cand_cnt[2 * s0] : 0u, cand_cnt[2 * s1 + 1] ? cand_cnt[2 * s0 + 1] : 0u,
			target_cnt[s0] ? cand_cnt[2 * s1] : 0u };
		for (int t = 0; t < 3; ++t)
		{
			off[t * (nslots + 1) + j] = run[t];
			run[t] += n[t];
		}
	}
}

// number of candidate faces of each (contour, role): needed only as "any"
// Whether (contour, role) s has candidate faces, and the candidate faces as a
// bitmap fcand[s * fwords + f / 64] bit f % 64.
__global__ void k_prp_any_faces(uint32_t F, uint32_t nslots2, const uint32_t* __restrict__ kring, uint32_t V,
	const uint32_t* __restrict__ hop_all, const uint8_t* __restrict__ own_all, const uint32_t* __restrict__ faces,
	uint32_t* __restrict__ target_cnt, uint64_t* __restrict__ fcand, uint32_t fwords)
{
	const uint32_t s = blockIdx.y;
	const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nslots2)
		return;
	const uint32_t* H = hop_all + (size_t)s * V;
	const uint8_t*	O = own_all + (size_t)s * V;
	const bool		in = f < F && cand_face(H, O, kring[s], faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]);
	const uint64_t	m = __ballot(in);
	if (__lane_id() == 0)
	{
		if (m != 0ull)
			atomicOr(&target_cnt[s], 1u);
		if ((f >> 6) < fwords)
			fcand[(size_t)s * fwords + (f >> 6)] = m;
	}
}


// ------------------------------------------------------------------ rays (cached_bvh.cpp:131-555)
constexpr uint32_t kHitLists = 64;
constexpr uint32_t kListStride = 16;
struct RayArgs
{
	uint32_t		V, E, nslots;
	const PrpSlot*	slots;
	const uint32_t* off; // [3][nslots + 1]
	const uint32_t* cand;
	const uint32_t* hop_all;
	const uint8_t*	own_all;
	const uint32_t* kring;
	const float*	x;
	const uint32_t* edges;
	const uint32_t* faces;
	PrpHit*			hits;		// kHitLists sub-lists of sub_cap entries
	uint32_t*		list_count; // counter of sub-list k at [kListStride * k]
	uint32_t		sub_cap;
	uint32_t*		status;
	uint32_t*		spill;
	const BvhNode*	nodes;
	const float*	tbox; // [12 * nslots] candidate edge boxes of role 1, then role 0
	const uint64_t* ecand; // candidate edge bitmaps per (contour, role), ewords each
	const uint64_t* fcand; // candidate face bitmaps per (contour, role), fwords each
	uint32_t		ewords, fwords;
	unsigned long long* tstat; // [3 kinds][4]: node visits, leaf tests, hits, candidate leaf tests; [12 + kind]: wave steps
};

__device__ __forceinline__ uint32_t slot_of(const uint32_t* off, uint32_t nslots, uint32_t r)
{
	uint32_t lo = 0, hi = nslots; // last j with off[j] <= r
	while (hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (off[mid] <= r)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// all 64 lanes call this; `acc` (meaningful in lane 0) counts the wave's hits
// Appends to the sub-list of the calling wave (one of kHitLists, by global wave
// index): a single shared counter would serialise every wave's append.
__device__ __forceinline__ void emit_hit(const RayArgs& a, bool ok, const PrpHit& h, unsigned long long& acc)
{
	const uint32_t k = (blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) % kHitLists;
	const uint32_t i = wave_append(&a.list_count[kListStride * k], ok);
	if (ok && i < a.sub_cap)
		a.hits[(size_t)k * a.sub_cap + i] = h;
	acc += (unsigned long long)__popcll(__ballot(ok));
}

// one atomic per wave for the counters a traversal accumulated in lane 0
__device__ __forceinline__ void flush_stat(unsigned long long* p, unsigned long long v)
{
	if (__lane_id() == 0 && v)
		atomicAdd(p, v);
}

// Per-lane request state shared by the three ray kinds.
struct RayReq
{
	uint32_t slot, src, contour;
	f3		 o, dir;
	float	 tmax; // ray length along dir
};

// VF (kind 0) and FV (kind 2): segment from a vertex along +dir (VF) or -dir (FV), length 1,
// against candidate faces of the other role (LineIntersection, t in [0, 1], incident faces skipped).
template <int Kind>
__global__ __launch_bounds__(kTravThreads, kRayBlocks / 256) void k_prp_ray_vert(RayArgs a)
{
	__shared__ TravSharedT<kRayStackLds> sh;
	const uint32_t		total = a.off[Kind * (a.nslots + 1) + a.nslots];
	const uint32_t*		off = a.off + Kind * (a.nslots + 1);
	const uint32_t		spill_slot = blockIdx.x * blockDim.x + threadIdx.x;
	// wave totals over the whole persistent loop, flushed once
	TravCount		   tc;
	unsigned long long hit_acc = 0; // lane 0: hits of this wave
	for (uint32_t base = xcd_block() * blockDim.x; base < total; base += gridDim.x * blockDim.x)
	{
		const uint32_t	   r = base + threadIdx.x;
		const bool		   active = r < total;
		RayReq			   q{};
		box3			   cb = box_empty();
		if (active)
		{
			q.slot = slot_of(off, a.nslots, r);
			const PrpSlot& sl = a.slots[q.slot];
			q.contour = sl.contour;
			const uint32_t s_src = 2 * sl.contour + (Kind == 0 ? 0u : 1u);
			q.src = a.cand[(size_t)s_src * (a.V + a.E) + (r - off[q.slot])];
			q.dir = mk3(sl.dir[0], sl.dir[1], sl.dir[2]);
			q.o = ld3(a.x, q.src);
			q.tmax = 1.0f;
			// the candidate target faces (role 1 for VF rays, role 0 for FV rays) lie in this box
			const float* tb = a.tbox + 12 * (size_t)q.slot + (Kind == 0 ? 0 : 6);
			cb = box3{ mk3(tb[0], tb[1], tb[2]), mk3(tb[3], tb[4], tb[5]) };
		}
		const f3 cast = Kind == 0 ? q.dir : -q.dir;
		const f3 inv = mk3(1.0f / cast.x, 1.0f / cast.y, 1.0f / cast.z);
		auto	 test = [&](const box3& b)
		{
			// slab test of cached_bvh.cpp:147-165 with a small outward tolerance
			const float t1x = (b.lo.x - q.o.x) * inv.x, t2x = (b.hi.x - q.o.x) * inv.x;
			const float t1y = (b.lo.y - q.o.y) * inv.y, t2y = (b.hi.y - q.o.y) * inv.y;
			const float t1z = (b.lo.z - q.o.z) * inv.z, t2z = (b.hi.z - q.o.z) * inv.z;
			float		tmin = fminf(t1x, t2x), tmx = fmaxf(t1x, t2x);
			tmin = fmaxf(tmin, fminf(t1y, t2y));
			tmx = fminf(tmx, fmaxf(t1y, t2y));
			tmin = fmaxf(tmin, fminf(t1z, t2z));
			tmx = fminf(tmx, fmaxf(t1z, t2z));
			const bool in_cb = (cb.lo.x <= b.hi.x) & (cb.hi.x >= b.lo.x) & (cb.lo.y <= b.hi.y)
				& (cb.hi.y >= b.lo.y) & (cb.lo.z <= b.hi.z) & (cb.hi.z >= b.lo.z);
			return in_cb && tmx + 1e-5f >= fmaxf(tmin, 0.0f) && tmin <= q.tmax + 1e-5f;
		};
		auto narrow = [&](bool valid, uint32_t rr, uint32_t f)
		{
			// staged candidates carry their request id rr, which belongs to lane rr - base
			// of this wave: its slot, contour, source vertex, direction and origin come
			// from that lane's registers (every lane shuffles, valid or not)
			const uint32_t src = (rr - base) & (kWave - 1);
			const uint32_t slot = __shfl(q.slot, src, kWave), contour = __shfl(q.contour, src, kWave);
			const uint32_t v = __shfl(q.src, src, kWave);
			const f3	   d = mk3(__shfl(q.dir.x, src, kWave), __shfl(q.dir.y, src, kWave), __shfl(q.dir.z, src, kWave));
			const f3	   o = mk3(__shfl(q.o.x, src, kWave), __shfl(q.o.y, src, kWave), __shfl(q.o.z, src, kWave));
			bool		   ok = false;
			PrpHit		   h{};
			if (valid)
			{
				const uint32_t s_dst = 2 * contour + (Kind == 0 ? 1u : 0u);
				const bool		is_cand = (a.fcand[(size_t)s_dst * a.fwords + (f >> 6)] >> (f & 63)) & 1ull;
				const uint32_t t0 = is_cand ? a.faces[3 * f] : v, t1 = is_cand ? a.faces[3 * f + 1] : v,
							   t2 = is_cand ? a.faces[3 * f + 2] : v;
				if (is_cand && v != t0 && v != t1 && v != t2)
				{
					const f3 e = Kind == 0 ? o + d * 1.0f : o + (-d) * 1.0f;
					f3		 bary;
					float	 t;
					if (line_intersection(o, e, ld3(a.x, t0), ld3(a.x, t1), ld3(a.x, t2), &bary, &t) && t >= 0.0f
						&& t <= 1.0f)
					{
						ok = true;
						h.slot = slot;
						h.type = Kind == 0 ? 0u : 2u;
						h.src = Kind == 0 ? v : f;
						h.dst = Kind == 0 ? f : v;
						h.b0 = bary.x;
						h.b1 = bary.y;
						h.dist = t * 1.0f;
						h.root = 0;
					}
				}
			}
			emit_hit(a, ok, h, hit_acc);
		};
		traverse_staged_test(sh, active, r, spill_slot, test, false, a.nodes, a.status, a.spill, 1u << 22, narrow,
			&tc);
	}
	flush_stat(&a.tstat[4 * Kind], tc.visits);
	flush_stat(&a.tstat[4 * Kind + 1], tc.tests);
	flush_stat(&a.tstat[4 * Kind + 2], hit_acc);
	flush_stat(&a.tstat[12 + Kind], tc.steps);
}

// EE (kind 1): source edge swept along dir by ray_max (h:597-613) against candidate edges of role 1
// (plane test, |det| >= 1e-10, Cramer solve with s in [0, ray_max], t1, t2 in [0, 1], shared vertices skipped).
// Beam condition of an EE sweep hit: with the sweep plane's unit normal pn and
// w = normalize(pn x dir) the in-plane axis across the sweep, the target
// endpoints' w . q range must meet [min, max] of w . p0, w . p1 widened by
// kEeBeamMargin.  A crossing at X = p0 + t e1 + s dir (t in [0, 1]) with the
// target segment has w . X in that range, so the condition only removes hits
// whose FP32 Cramer solve is so ill-conditioned (target nearly in the sweep
// plane) that the computed crossing lies more than the margin off the swept
// strip.  The oracle applies the same test.
constexpr float kEeBeamMargin = 1e-4f;
// rounding margin of the EE sweep's node culls (FP32 dot products of ~1 m coordinates)
constexpr float kEeNodeMargin = 1e-4f;
// The beam condition with w = normalize(pn x dir) given (each request computes it
// once); w = 0 stands for |pn x dir| < 1e-10, where the condition keeps the edge.
__device__ __forceinline__ bool ee_in_beam(f3 wn, f3 p0, f3 p1, f3 q0, f3 q1)
{
	if (wn.x == 0.0f && wn.y == 0.0f && wn.z == 0.0f)
		return true;
	const float w0 = dot3(wn, p0), w1 = dot3(wn, p1), u0 = dot3(wn, q0), u1 = dot3(wn, q1);
	return fmaxf(u0, u1) >= fminf(w0, w1) - kEeBeamMargin && fminf(u0, u1) <= fmaxf(w0, w1) + kEeBeamMargin;
}

__global__ __launch_bounds__(kTravThreads, kEeWavesPerSimd) void k_prp_ray_edge(RayArgs a)
{
	__shared__ TravSharedT<kRayStackLds> sh;
	const uint32_t		  total = a.off[1 * (a.nslots + 1) + a.nslots];
	const uint32_t*		  off = a.off + (a.nslots + 1);
	const uint32_t		  spill_slot = blockIdx.x * blockDim.x + threadIdx.x;
	// wave totals over the whole persistent loop, flushed once
	TravCount		   tc;
	unsigned long long hit_acc = 0, cand_acc = 0; // lane 0: hits and candidate leaf tests of the wave
	for (uint32_t base = xcd_block() * blockDim.x; base < total; base += gridDim.x * blockDim.x)
	{
		const uint32_t r = base + threadIdx.x;
		bool		   active = r < total;
		// per-request node-cull constants (see the test below); slot, e, rmax, pn, pd and
		// wn also serve the leaf tests of this request, which read them from this lane
		box3	 ib = box_empty();
		f3		 pn = mk3(0.0f, 0.0f, 0.0f), wn = mk3(0.0f, 0.0f, 0.0f);
		float	 slab_lo = -1e30f, slab_hi = 1e30f, beam_lo = -1e30f, beam_hi = 1e30f;
		uint32_t slot = 0, e = 0;
		float	 rmax = 1.0f, pd = 0.0f;
		if (active)
		{
			slot = slot_of(off, a.nslots, r);
			const PrpSlot& sl = a.slots[slot];
			e = a.cand[(size_t)(2 * sl.contour) * (a.V + a.E) + a.V + (r - off[slot])];
			const f3	   dir = mk3(sl.dir[0], sl.dir[1], sl.dir[2]);
			const f3	   p0 = ld3(a.x, a.edges[2 * e]), p1 = ld3(a.x, a.edges[2 * e + 1]);
			const float	   dl2 = dot3(dir, dir);
			if (dl2 > 1e-12f)
			{
				const float smin = fminf(dot3(p0, dir), dot3(p1, dir));
				rmax = fminf(fmaxf((sl.ee_pmax - smin) / dl2 + 1e-5f, 0.0f), 1.0f);
			}
			if (rmax <= 0.0f)
				active = false;
			box3 sw = box_add(box_add(box_add(box_of(p0), p1), p0 + dir * rmax), p1 + dir * rmax);
			sw = box_inflate(sw, 1e-5f);
			const float* tb = a.tbox + 12 * (size_t)slot;
			// every candidate target edge lies in the candidate box, so a kept edge meets the swept
			// box inside it: one test against the intersection replaces the two box tests
			ib.lo = mk3(fmaxf(sw.lo.x, tb[0]), fmaxf(sw.lo.y, tb[1]), fmaxf(sw.lo.z, tb[2]));
			ib.hi = mk3(fminf(sw.hi.x, tb[3]), fminf(sw.hi.y, tb[4]), fminf(sw.hi.z, tb[5]));
			if (ib.lo.x > ib.hi.x || ib.lo.y > ib.hi.y || ib.lo.z > ib.hi.z)
				active = false; // no candidate edge meets the sweep
			// the leaf's plane prune (same normal): a kept target edge has a point within 1e-6 of the plane
			pn = cross3(p1 - p0, dir);
			const float pl = sqrtf(dot3(pn, pn));
			if (pl >= 1e-10f)
			{
				pn = pn * (1.0f / pl);
				pd = dot3(pn, p0);
				slab_lo = pd - (1e-6f + kEeNodeMargin);
				slab_hi = pd + (1e-6f + kEeNodeMargin);
				// in-plane axis across the sweep: w . dir = 0, so the swept parallelogram
				// p0 + t e1 + s dir covers only [min, max] of w . p0, w . p1 along it
				wn = cross3(pn, dir);
				const float wl = sqrtf(dot3(wn, wn));
				if (wl >= 1e-10f)
				{
					wn = wn * (1.0f / wl);
					const float w0 = dot3(wn, p0), w1 = dot3(wn, p1);
					beam_lo = fminf(w0, w1) - (kEeBeamMargin + kEeNodeMargin);
					beam_hi = fmaxf(w0, w1) + (kEeBeamMargin + kEeNodeMargin);
				}
				else
					wn = mk3(0.0f, 0.0f, 0.0f);
			}
			else
				pn = mk3(0.0f, 0.0f, 0.0f); // no plane prune at the leaf either
		}
		// Node culling.  A node can hold a kept target edge only if its box meets (1) the swept box
		// and (2) the box of the slot's candidate target edges -- tested together as their
		// intersection ib, since every kept edge is a candidate and lies in (2) -- and (3) the slab
		// |n.x - pd| <= 1e-6 that the leaf's plane prune requires of one point of the edge.
		// (4) The leaf keeps only an edge whose endpoints' w . q range meets [min, max] of
		// w . p0, w . p1 widened by kEeBeamMargin, w the in-plane axis across the sweep
		// (ee_in_beam; a crossing at X = p0 + t e1 + s dir with t in [0, 1] has w . X in that range).
		// (1)-(4) hold for the leaf box of any kept edge and so for its ancestors.  (3) and (4) test
		// the range of n . x (w . x) over the box, whose extremes lie at the corners picked by the
		// signs of n (w); kEeNodeMargin covers the FP32 rounding of the node and leaf dot products
		// (~1e-6 here), so the culls stay conservative.
		auto test = [&](const box3& b)
		{
			const bool	in_box = (ib.lo.x <= b.hi.x) & (ib.hi.x >= b.lo.x) & (ib.lo.y <= b.hi.y)
				& (ib.hi.y >= b.lo.y) & (ib.lo.z <= b.hi.z) & (ib.hi.z >= b.lo.z);
			const float pmax = fmaf(pn.x, pn.x >= 0.0f ? b.hi.x : b.lo.x,
				fmaf(pn.y, pn.y >= 0.0f ? b.hi.y : b.lo.y, pn.z * (pn.z >= 0.0f ? b.hi.z : b.lo.z)));
			const float pmin = fmaf(pn.x, pn.x >= 0.0f ? b.lo.x : b.hi.x,
				fmaf(pn.y, pn.y >= 0.0f ? b.lo.y : b.hi.y, pn.z * (pn.z >= 0.0f ? b.lo.z : b.hi.z)));
			const float wmax = fmaf(wn.x, wn.x >= 0.0f ? b.hi.x : b.lo.x,
				fmaf(wn.y, wn.y >= 0.0f ? b.hi.y : b.lo.y, wn.z * (wn.z >= 0.0f ? b.hi.z : b.lo.z)));
			const float wmin = fmaf(wn.x, wn.x >= 0.0f ? b.lo.x : b.hi.x,
				fmaf(wn.y, wn.y >= 0.0f ? b.lo.y : b.hi.y, wn.z * (wn.z >= 0.0f ? b.lo.z : b.hi.z)));
			return in_box & (pmax >= slab_lo) & (pmin <= slab_hi) & (wmax >= beam_lo) & (wmin <= beam_hi);
		};
		auto narrow = [&](bool valid, uint32_t rr, uint32_t g)
		{
			// The staged candidate belongs to request rr, made by lane rr - base of this
			// wave (all 64 lanes are here): its slot, source edge, sweep length, plane
			// (pn, pd) and beam axis wn come from that lane, bit-identical to a
			// recomputation from the slot and the edge (same expressions, same inputs).
			const uint32_t src = (rr - base) & (kWave - 1);
			const uint32_t qs = __shfl(slot, src, kWave), qe = __shfl(e, src, kWave);
			const float	   q_rmax = __shfl(rmax, src, kWave), q_pd = __shfl(pd, src, kWave);
			const f3	   q_pn = mk3(__shfl(pn.x, src, kWave), __shfl(pn.y, src, kWave), __shfl(pn.z, src, kWave));
			const f3	   q_wn = mk3(__shfl(wn.x, src, kWave), __shfl(wn.y, src, kWave), __shfl(wn.z, src, kWave));
			bool		   ok = false;
			PrpHit		   h{};
			if (valid)
			{
				const PrpSlot&	sl = a.slots[qs];
				const uint32_t	s1 = 2 * sl.contour + 1;
				const bool		is_cand = (a.ecand[(size_t)s1 * a.ewords + (g >> 6)] >> (g & 63)) & 1ull;
				uint32_t		a0 = 0, a1 = 0, b0 = 1, b1 = 1;
				if (is_cand)
				{
					a0 = a.edges[2 * qe], a1 = a.edges[2 * qe + 1];
					b0 = a.edges[2 * g], b1 = a.edges[2 * g + 1];
				}
				const bool cand = is_cand && a0 != b0 && a0 != b1 && a1 != b0 && a1 != b1;
				cand_acc += (unsigned long long)__popcll(__ballot(cand)); // lane 0 is valid whenever narrow runs
				if (cand)
				{
					const f3 dir = mk3(sl.dir[0], sl.dir[1], sl.dir[2]);
					const f3 p0 = ld3(a.x, a0), p1 = ld3(a.x, a1), e1 = p1 - p0;
					const f3 q0 = ld3(a.x, b0), q1 = ld3(a.x, b1);
					// pn = 0 when |e1 x dir| < 1e-10 (no plane prune)
					const bool use_plane = q_pn.x != 0.0f || q_pn.y != 0.0f || q_pn.z != 0.0f;
					bool	   keep = true;
					if (use_plane)
					{
						const float d0 = dot3(q_pn, q0) - q_pd, d1 = dot3(q_pn, q1) - q_pd;
						keep = !((d0 > 1e-6f && d1 > 1e-6f) || (d0 < -1e-6f && d1 < -1e-6f))
							&& ee_in_beam(q_wn, p0, p1, q0, q1);
					}
					if (keep)
					{
						const f3	e2 = q1 - q0, rv = q0 - p0, nn = cross3(e1, e2);
						const float det = -dot3(dir, nn);
						if (fabsf(det) >= 1e-10f)
						{
							const float inv = 1.0f / det;
							const float s_ = -dot3(rv, cross3(e1, e2)) * inv;
							const float t1 = -dot3(dir, cross3(rv, e2)) * inv;
							const float t2 = dot3(dir, cross3(e1, rv)) * inv;
							if (s_ >= 0.0f && s_ <= q_rmax && t1 >= 0.0f && t1 <= 1.0f && t2 >= 0.0f && t2 <= 1.0f)
							{
								ok = true;
								h.slot = qs, h.type = 1u, h.src = qe, h.dst = g;
								h.b0 = 1.0f - t1, h.b1 = 1.0f - t2, h.dist = s_, h.root = 0;
							}
						}
					}
				}
			}
			emit_hit(a, ok, h, hit_acc);
		};
		traverse_staged_test(sh, active, r, spill_slot, test, false, a.nodes, a.status, a.spill, 1u << 22, narrow,
			&tc);
	}
	flush_stat(&a.tstat[4], tc.visits);
	flush_stat(&a.tstat[4 + 1], tc.tests);
	flush_stat(&a.tstat[4 + 2], hit_acc);
	flush_stat(&a.tstat[4 + 3], cand_acc);
	flush_stat(&a.tstat[12 + 1], tc.steps);
}

// ------------------------------------------------------------------ hit graph
struct HitArgs
{
	uint32_t		n, V, prim_bits;
	const PrpHit*	h;		 // sorted by (slot, type, src)
	const uint32_t* keys;	 // sorted keys
	const uint32_t* dst;	 // target primitive of each sorted hit
	const uint32_t* range;	 // [2 * slots] hit range of each slot
	const uint64_t* table;	 // bucket key -> first index (generation | key | index), or null
	uint32_t		table_cap;
	uint64_t		gen;
	uint32_t*		parent;	 // union-find
	const uint32_t *vv_off, *vv, *ve_off, *ve, *vf_off, *vf;
	const uint32_t *edges, *faces, *fae, *eaf;
};

__device__ __forceinline__ uint32_t hit_key(uint32_t slot, uint32_t type, uint32_t prim, uint32_t pb)
{
	return (slot << (pb + 2)) | (type << pb) | prim;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* k, uint32_t n, uint32_t v)
{
	uint32_t lo = 0, hi = n;
	while (lo < hi)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (k[mid] < v)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

// Lock-free union-find: roots are linked by CAS (larger index under smaller);
// find halves the path.  A halving store only ever writes to a non-root
// element (never a CAS target) and writes one of its ancestors, so every
// parent chain stays valid under concurrent unions.
__device__ __forceinline__ uint32_t uf_find(uint32_t* p, uint32_t x)
{
	for (;;)
	{
		const uint32_t q = __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (q == x)
			return x;
		const uint32_t g = __hip_atomic_load(&p[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (g != q)
			__hip_atomic_store(&p[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		x = g;
	}
}

// Union for the hit graph: finds read parents with cheap (workgroup-scope,
// possibly stale) loads, which still return ancestors, since parents only move
// up; the link is a device-scope CAS, and a failed CAS returns the element's
// current parent, from which the walk resumes.  Each failure replaces the
// larger of the two candidates by a smaller index, so the loop ends.
__device__ __forceinline__ uint32_t uf_load_cached(const uint32_t* p, uint32_t x)
{
	return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t uf_find_cached(const uint32_t* p, uint32_t x)
{
	for (;;)
	{
		const uint32_t q = uf_load_cached(p, x);
		if (q == x)
			return x;
		x = q;
	}
}

// Returns the root the two sets share after the union (an ancestor of a and of
// b); a caller that unions one element with many passes it back as `a`, so
// each later find starts near the root instead of at the element.
__device__ __forceinline__ uint32_t uf_union_cached(uint32_t* p, uint32_t a, uint32_t b)
{
	a = uf_find_cached(p, a);
	b = uf_find_cached(p, b);
	while (a != b)
	{
		const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
		const uint32_t old = atomicCAS(&p[hi], hi, lo);
		if (old == hi)
			return lo;
		// hi was linked meanwhile: continue from its current parent (< hi)
		const uint32_t up = uf_find_cached(p, old);
		if (a == hi)
			a = up;
		else
			b = up;
	}
	return a;
}

__device__ __forceinline__ void uf_union(uint32_t* p, uint32_t a, uint32_t b)
{
	for (;;)
	{
		a = uf_find(p, a);
		b = uf_find(p, b);
		if (a == b)
			return;
		if (a < b)
		{
			const uint32_t t = a;
			a = b;
			b = t;
		}
		if (atomicCAS(&p[a], a, b) == a)
			return;
	}
}

__device__ __forceinline__ bool face_has_vert(const uint32_t* faces, uint32_t f, uint32_t v)
{
	return faces[3 * f] == v || faces[3 * f + 1] == v || faces[3 * f + 2] == v;
}
__device__ __forceinline__ bool face_has_edge(const uint32_t* eaf, uint32_t f, uint32_t e)
{
	return eaf[2 * e] == f || eaf[2 * e + 1] == f;
}
__device__ __forceinline__ bool edge_has_vert(const uint32_t* edges, uint32_t e, uint32_t v)
{
	return edges[2 * e] == v || edges[2 * e + 1] == v;
}
__device__ __forceinline__ bool verts_adjacent(const HitArgs& a, uint32_t u, uint32_t v)
{
	for (uint32_t k = a.vv_off[u]; k < a.vv_off[u + 1]; ++k)
		if (a.vv[k] == v)
			return true;
	return false;
}
// e2 is one of the other two edges of a face adjacent to e1 (edge_adj_edges_ext, init_mesh_data.cpp:505-531)
__device__ __forceinline__ bool edges_adjacent(const HitArgs& a, uint32_t e1, uint32_t e2)
{
	if (e1 == e2)
		return false;
	for (int k = 0; k < 2; ++k)
	{
		const uint32_t f = a.eaf[2 * e1 + k];
		if (f != kNone && (a.fae[3 * f] == e2 || a.fae[3 * f + 1] == e2 || a.fae[3 * f + 2] == e2))
			return true;
	}
	return false;
}

// Unions hit i's set (through ri, an ancestor of i) with the sets of the hits
// j[q] (kNone entries are skipped) and returns the shared root.  The parent
// chains of the j are walked side by side first, so their loads overlap.
template <int N>
__device__ __forceinline__ uint32_t uf_union_n(uint32_t* p, uint32_t i, uint32_t ri, const uint32_t (&j)[N])
{
	uint32_t r[N], live = 0;
#pragma unroll
	for (int q = 0; q < N; ++q)
	{
		r[q] = j[q];
		if (j[q] != kNone && j[q] != i)
			live |= 1u << q;
	}
	const uint32_t linked = live;
	while (live)
	{
		uint32_t up[N];
#pragma unroll
		for (int q = 0; q < N; ++q)
			up[q] = (live >> q) & 1u ? uf_load_cached(p, r[q]) : r[q];
#pragma unroll
		for (int q = 0; q < N; ++q)
		{
			if (up[q] == r[q])
				live &= ~(1u << q);
			r[q] = up[q];
		}
	}
#pragma unroll
	for (int q = 0; q < N; ++q)
		if ((linked >> q) & 1u)
			ri = uf_union_cached(p, ri, r[q]);
	return ri;
}

__device__ __forceinline__ uint64_t mix64(uint64_t k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return k;
}

// Bucket table: entry = generation (8 bits) | key (32 bits) | first index (24 bits);
// entries of other generations are empty.  One insert per bucket start.
__global__ void k_prp_bucket_insert(uint32_t n, const uint32_t* __restrict__ keys, uint64_t* table, uint32_t cap,
	uint64_t gen)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || (i > 0 && keys[i] == keys[i - 1]))
		return;
	const uint64_t e = (gen << 56) | ((uint64_t)keys[i] << 24) | i;
	for (uint32_t s = (uint32_t)mix64(keys[i]) & (cap - 1), c = 0; c < cap; s = (s + 1) & (cap - 1), ++c)
	{
		uint64_t cur = __hip_atomic_load(&table[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		while ((cur >> 56) != gen)
		{
			const uint64_t prev =
				atomicCAS((unsigned long long*)&table[s], (unsigned long long)cur, (unsigned long long)e);
			if (prev == cur)
				return;
			cur = prev;
		}
	}
}

__device__ __forceinline__ uint32_t bucket_first(const HitArgs& a, uint32_t key)
{
	for (uint32_t s = (uint32_t)mix64(key) & (a.table_cap - 1), c = 0; c < a.table_cap;
		s = (s + 1) & (a.table_cap - 1), ++c)
	{
		const uint64_t e = a.table[s];
		if ((e >> 56) != a.gen)
			return kNone;
		if ((uint32_t)(e >> 24) == key)
			return (uint32_t)(e & 0xFFFFFFu);
	}
	return kNone;
}

// First sorted index of the bucket (slot, type, source prim), or kNone: through
// the table, or a search of the slot's range.
__device__ __forceinline__ uint32_t bucket_start(const HitArgs& a, uint32_t slot, uint32_t key)
{
	if (a.table)
		return bucket_first(a, key);
	const uint32_t l0 = a.range[2 * slot], hi = a.range[2 * slot + 1];
	const uint32_t lo = l0 + lower_bound_u32(a.keys + l0, hi - l0, key);
	return (lo < hi && a.keys[lo] == key) ? lo : kNone;
}

// Index of the hit of bucket `key` (starting at or before `lo`) with dst d, or
// kNone.  Hits are sorted by (key, dst) and a (source, dst) pair occurs at most
// once per slot: d is found by galloping from lo, so large buckets cost
// O(log size).  *next gets the first index with dst >= d, where a search for a
// larger d of the same bucket may start.
__device__ __forceinline__ uint32_t gallop_dst(const HitArgs& a, uint32_t key, uint32_t lo, uint32_t hi, uint32_t d,
	uint32_t* next)
{
	auto	 less = [&](uint32_t m) { return a.keys[m] == key && a.dst[m] < d; };
	uint32_t top = hi;
	for (uint32_t step = 1;; step <<= 1)
	{
		const uint32_t m = lo + step - 1;
		if (m >= hi || !less(m))
		{
			top = m < hi ? m : hi;
			break;
		}
		lo = m + 1;
	}
	while (lo < top)
	{
		const uint32_t mid = (lo + top) >> 1;
		if (less(mid))
			lo = mid + 1;
		else
			top = mid;
	}
	*next = lo;
	return (lo < hi && a.keys[lo] == key && a.dst[lo] == d) ? lo : kNone;
}

// find_adjacent_hits (intersection_resolver_helper.cpp:771-902): every rule
// relates hit i to hits of a neighbouring source primitive whose target lies in
// a small set fixed by i's target, so each rule is a few exact lookups.  The
// relation is symmetric (VF-FV, VF-EE and FV-EE are one condition seen from
// either side, and the same-type rules mirror), so each adjacent pair is looked
// up from one side only: cross-type pairs from the VF / FV hit, same-type pairs
// from the hit whose varying primitive is smaller.
__global__ void k_prp_hit_union(HitArgs a)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n)
		return;
	const PrpHit   L = a.h[i];
	const uint32_t s = L.slot, hi = a.range[2 * s + 1];
	uint32_t	   ri = i; // an ancestor of i: where the next union's find of i starts
	// the hits (type, prim) -> d for targets in ascending order: one bucket lookup,
	// then galloping searches that resume where the previous one ended.  A search
	// in hit i's own bucket for targets above i's target starts at i + 1 instead
	// (the bucket is sorted by target), with no table probe.
	auto link_list = [&](uint32_t type, uint32_t prim, const uint32_t* __restrict__ d, uint32_t n,
						 uint32_t from = kNone)
	{
		const uint32_t key = hit_key(s, type, prim, a.prim_bits);
		uint32_t	   lo = from != kNone ? from : bucket_start(a, s, key);
		if (lo == kNone || lo >= hi)
			return;
		for (uint32_t q = 0; q < n; ++q)
		{
			const uint32_t j = gallop_dst(a, key, lo, hi, d[q], &lo);
			if (j != kNone && j != i)
				ri = uf_union_cached(a.parent, ri, j);
		}
	};
	// the same for up to four targets held in registers (kNone entries, last, are skipped);
	// the unions follow the four searches, so their parent walks overlap
	auto link4 = [&](uint32_t type, uint32_t prim, uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3,
					 uint32_t from = kNone)
	{
		const uint32_t key = hit_key(s, type, prim, a.prim_bits);
		uint32_t	   lo = from != kNone ? from : bucket_start(a, s, key);
		if (lo == kNone || lo >= hi)
			return;
		const uint32_t d[4] = { d0, d1, d2, d3 };
		uint32_t	   j[4] = { kNone, kNone, kNone, kNone };
#pragma unroll
		for (int q = 0; q < 4; ++q)
			if (d[q] != kNone)
				j[q] = gallop_dst(a, key, lo, hi, d[q], &lo);
		ri = uf_union_n(a.parent, i, ri, j);
	};
	auto cswap = [](uint32_t& x, uint32_t& y)
	{
		const uint32_t lo = min(x, y), hi2 = max(x, y);
		x = lo, y = hi2;
	};
	if (L.type == 0) // VF(v, f)
	{
		const uint32_t v = L.src, f = L.dst;
		// VF(w, f), w adjacent to v (w > v: the pair is symmetric)
		for (uint32_t k = a.vv_off[v]; k < a.vv_off[v + 1]; ++k)
			if (a.vv[k] > v)
				link4(0, a.vv[k], f, kNone, kNone, kNone);
		// FV(f', u), f' a face of v, u a vertex of f
		uint32_t u0 = a.faces[3 * f], u1 = a.faces[3 * f + 1], u2 = a.faces[3 * f + 2];
		cswap(u0, u1), cswap(u1, u2), cswap(u0, u1);
		for (uint32_t k = a.vf_off[v]; k < a.vf_off[v + 1]; ++k)
			link4(2, a.vf[k], u0, u1, u2, kNone);
		// EE(e, e'), e an edge of v, e' an edge of f
		uint32_t g0 = a.fae[3 * f], g1 = a.fae[3 * f + 1], g2 = a.fae[3 * f + 2];
		cswap(g0, g1), cswap(g1, g2), cswap(g0, g1);
		for (uint32_t k = a.ve_off[v]; k < a.ve_off[v + 1]; ++k)
			link4(1, a.ve[k], g0, g1, g2, kNone);
	}
	else if (L.type == 2) // FV(f, v)
	{
		const uint32_t f = L.src, v = L.dst;
		// (VF(u, f'), u a vertex of f, f' a face of v: found from the VF side)
		// FV(f, u), u adjacent to v (u > v): the neighbours of v are listed ascending;
		// i's own bucket, after i
		const uint32_t q0 = a.vv_off[v], q1 = a.vv_off[v + 1];
		uint32_t	   q = q0;
		while (q < q1 && a.vv[q] <= v)
			++q;
		if (q < q1)
			link_list(2, f, a.vv + q, q1 - q, i + 1);
		// EE(e, e'), e an edge of f, e' an edge of v (listed ascending)
		for (int k = 0; k < 3; ++k)
			link_list(1, a.fae[3 * f + k], a.ve + a.ve_off[v], a.ve_off[v + 1] - a.ve_off[v]);
	}
	else // EE(e1, e2)
	{
		const uint32_t e1 = L.src, e2 = L.dst;
		// (VF and FV partners are found from their side)
		// EE(e1, e'), e' another edge of a face of e2 (e' > e2), ascending
		uint32_t t[4] = { kNone, kNone, kNone, kNone };
#pragma unroll
		for (int q2 = 0; q2 < 2; ++q2)
		{
			const uint32_t g = a.eaf[2 * e2 + q2];
			if (g == kNone)
				continue;
			const uint32_t x0 = a.fae[3 * g], x1 = a.fae[3 * g + 1], x2 = a.fae[3 * g + 2];
			// the two edges of g other than e2
			const uint32_t o0 = x0 == e2 ? x1 : x0, o1 = (x0 == e2 || x1 == e2) ? x2 : x1;
			t[2 * q2] = o0 > e2 ? o0 : kNone;
			t[2 * q2 + 1] = o1 > e2 ? o1 : kNone;
		}
		cswap(t[0], t[1]), cswap(t[2], t[3]), cswap(t[0], t[2]), cswap(t[1], t[3]), cswap(t[1], t[2]);
		link4(1, e1, t[0], t[1], t[2], t[3], i + 1); // i's own bucket, after i
		// EE(e', e2), e' another edge of a face of e1 (e' > e1); one union batch
		uint32_t j2[4] = { kNone, kNone, kNone, kNone };
#pragma unroll
		for (int k = 0; k < 2; ++k)
		{
			const uint32_t f = a.eaf[2 * e1 + k];
			if (f == kNone)
				continue;
			const uint32_t x0 = a.fae[3 * f], x1 = a.fae[3 * f + 1], x2 = a.fae[3 * f + 2];
			// the two edges of f other than e1
			const uint32_t o[2] = { x0 == e1 ? x1 : x0, (x0 == e1 || x1 == e1) ? x2 : x1 };
#pragma unroll
			for (int m = 0; m < 2; ++m)
				if (o[m] > e1)
				{
					const uint32_t key = hit_key(s, 1, o[m], a.prim_bits);
					uint32_t	   lo = bucket_start(a, s, key);
					if (lo != kNone)
						j2[2 * k + m] = gallop_dst(a, key, lo, hi, e2, &lo);
				}
		}
		ri = uf_union_n(a.parent, i, ri, j2);
	}
}

// ------------------------------------------------------------------ clusters (:3506-3685)
struct CompArgs
{
	uint32_t		n, nslots, V, c0;
	const PrpHit*	h;
	const uint32_t* parent;
	const PrpSlot*	slots;
	const uint32_t *flag_off, *flags, *edges, *faces;
	uint32_t*		attached;  // [n] per root
	uint32_t*		count;	   // [n] per root
	uint64_t*		key;	   // [n] per root, min canonical key
	uint32_t*		cov;	   // [n] per root
	uint32_t*		roots;	   // attached roots
	uint32_t*		nroots;
	uint64_t*		hash;
	uint32_t		hash_cap;
	uint64_t		gen; // support-key generation of this evaluation
	uint64_t*		slot_best; // [2 * nslots]: (cov << 32 | count), then key
	uint32_t*		slot_root; // [nslots]
	const uint32_t *cpair_off, *cpair_rec;
};

__device__ __forceinline__ uint32_t src_verts(const uint32_t* edges, const uint32_t* faces, const PrpHit& h,
	uint32_t* o)
{
	if (h.type == 0)
		return o[0] = h.src, 1u;
	if (h.type == 1)
		return o[0] = edges[2 * h.src], o[1] = edges[2 * h.src + 1], 2u;
	o[0] = faces[3 * h.src], o[1] = faces[3 * h.src + 1], o[2] = faces[3 * h.src + 2];
	return 3u;
}
__device__ __forceinline__ uint32_t dst_verts(const uint32_t* edges, const uint32_t* faces, const PrpHit& h,
	uint32_t* o)
{
	if (h.type == 0)
	{
		o[0] = faces[3 * h.dst], o[1] = faces[3 * h.dst + 1], o[2] = faces[3 * h.dst + 2];
		return 3u;
	}
	if (h.type == 1)
		return o[0] = edges[2 * h.dst], o[1] = edges[2 * h.dst + 1], 2u;
	return o[0] = h.dst, 1u;
}

__device__ __forceinline__ bool has_flag(const uint32_t* off, const uint32_t* flags, uint32_t v, uint32_t f)
{
	uint32_t lo = off[v], hi = off[v + 1];
	while (lo < hi)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (flags[mid] < f)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo < off[v + 1] && flags[lo] == f;
}

__device__ __forceinline__ uint64_t canonical_key(const PrpHit& h)
{
	return ((uint64_t)h.type << 62) | ((uint64_t)(h.src & 0x3fffffffu) << 31) | (uint64_t)(h.dst & 0x7fffffffu);
}

// after all unions: every element points at its root
// After the unions no root changes, so a walk over cached (possibly already
// shortened) parents reaches the same root.
__global__ void k_prp_uf_compress(uint32_t n, uint32_t* parent)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		parent[i] = uf_find_cached(parent, i);
}

__global__ void k_prp_uf_init(uint32_t n, uint32_t* parent, uint32_t* attached, uint32_t* count, uint64_t* key,
	uint32_t* cov)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	parent[i] = i, attached[i] = 0, count[i] = 0, key[i] = ~0ull, cov[i] = 0;
}

// root of every hit, boundary attachment (is_hit_adj_contour :378-412), size and canonical key per root
__global__ void k_prp_components(CompArgs a, PrpHit* hs)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	const bool	   valid = i < a.n;
	uint32_t	   r = 0;
	bool		   adj = false;
	uint64_t	   key = ~0ull;
	if (valid)
	{
		r = uf_find(const_cast<uint32_t*>(a.parent), i);
		PrpHit h = hs[i];
		h.root = r;
		hs[i] = h;
		const uint32_t c = a.c0 + a.slots[h.slot].contour;
		uint32_t	   sv[3], dv[3];
		const uint32_t ns = src_verts(a.edges, a.faces, h, sv), nd = dst_verts(a.edges, a.faces, h, dv);
		adj = true;
		for (uint32_t k = 0; k < ns; ++k)
			adj = adj && has_flag(a.flag_off, a.flags, sv[k], 2 * c);
		for (uint32_t k = 0; k < nd; ++k)
			adj = adj && has_flag(a.flag_off, a.flags, dv[k], 2 * c + 1);
		key = canonical_key(h);
	}
	// one set of atomics per distinct root in the wave (hits are sorted, so a wave
	// mostly holds one or two clusters, which would otherwise serialise on one address)
	uint64_t pending = __ballot(valid);
	while (pending)
	{
		const int	   leader = __ffsll((unsigned long long)pending) - 1;
		const uint32_t lr = __shfl(r, leader, kWave);
		const bool	   mine = ((pending >> __lane_id()) & 1ull) && r == lr;
		const uint64_t m = __ballot(mine);
		const bool	   any_adj = __ballot(mine && adj) != 0ull;
		uint64_t	   k = mine ? key : ~0ull;
		for (int o = 32; o > 0; o >>= 1)
		{
			const uint64_t t = __shfl_xor(k, o, kWave);
			k = t < k ? t : k;
		}
		if ((int)__lane_id() == leader)
		{
			if (any_adj)
				atomicOr(&a.attached[lr], 1u);
			atomicAdd(&a.count[lr], (uint32_t)__popcll(m));
			atomicMin((unsigned long long*)&a.key[lr], (unsigned long long)k);
		}
		pending &= ~m;
	}
}

__device__ __forceinline__ uint64_t hmix(uint64_t k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return k;
}

// Open-addressing set of 64-bit keys whose top 8 bits are the evaluation's
// generation: an entry of another generation (or the ~0 fill) counts as empty,
// so the table is cleared only when the generation counter wraps.
__device__ __forceinline__ uint64_t key_gen(uint64_t k) { return k >> 56; }

__device__ __forceinline__ void hash_insert(uint64_t* t, uint32_t cap, uint64_t k)
{
	const uint64_t gen = key_gen(k);
	for (uint32_t i = (uint32_t)hmix(k) & (cap - 1), n = 0; n < cap; i = (i + 1) & (cap - 1), ++n)
	{
		uint64_t cur = __hip_atomic_load(&t[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		for (;;)
		{
			if (cur == k)
				return;
			if (key_gen(cur) == gen)
				break; // another key of this evaluation: probe on
			const uint64_t prev =
				atomicCAS((unsigned long long*)&t[i], (unsigned long long)cur, (unsigned long long)k);
			if (prev == cur)
				return;
			cur = prev; // lost the race: look at what was written
		}
	}
}

__device__ __forceinline__ bool hash_has(const uint64_t* t, uint32_t cap, uint64_t k)
{
	for (uint32_t i = (uint32_t)hmix(k) & (cap - 1), n = 0; n < cap; i = (i + 1) & (cap - 1), ++n)
	{
		const uint64_t v = t[i];
		if (v == k)
			return true;
		if (key_gen(v) != key_gen(k))
			return false;
	}
	return false;
}

// generation (8 bits) | root hit (29 bits) | vertex (26 bits) | side (1 bit)
__device__ __forceinline__ uint64_t support_key(uint64_t gen, uint32_t root, uint32_t v, uint32_t side)
{
	return (gen << 56) | ((uint64_t)root << 27) | ((uint64_t)v << 1) | side;
}

// support vertices of every hit in an attached cluster: (root, vertex, 0 = source / 1 = target)
__global__ void k_prp_support(CompArgs a)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.n)
		return;
	const PrpHit h = a.h[i];
	if (!a.attached[h.root])
		return;
	if (h.root == i)
		a.roots[atomicAdd(a.nroots, 1u)] = i;
	uint32_t	   sv[3], dv[3];
	const uint32_t ns = src_verts(a.edges, a.faces, h, sv), nd = dst_verts(a.edges, a.faces, h, dv);
	for (uint32_t k = 0; k < ns; ++k)
		hash_insert(a.hash, a.hash_cap, support_key(a.gen, h.root, sv[k], 0));
	for (uint32_t k = 0; k < nd; ++k)
		hash_insert(a.hash, a.hash_cap, support_key(a.gen, h.root, dv[k], 1));
}

// EF-anchor coverage of each attached cluster (component_anchor_coverage :3587-3621):
// pair (edge in role m, face in role m^1) is covered when an edge vertex is a
// source vertex (m = 0) / target vertex (m = 1) of the cluster and a face vertex
// is on the other side.  One workgroup per cluster.
__global__ void k_prp_coverage(CompArgs a)
{
	const uint32_t nr = *a.nroots;
	for (uint32_t b = blockIdx.x; b < nr; b += gridDim.x)
	{
		const uint32_t r = a.roots[b];
		const PrpHit   h = a.h[r];
		const uint32_t c = a.c0 + a.slots[h.slot].contour;
		const uint32_t p0 = a.cpair_off[c], p1 = a.cpair_off[c + 1];
		uint32_t	   cnt = 0;
		for (uint32_t p = p0 + threadIdx.x; p < p1; p += blockDim.x)
		{
			const uint32_t* q = a.cpair_rec + 6 * p;
			const uint32_t	m = q[5];
			const uint32_t	es = m == 0 ? 0u : 1u, fs = es ^ 1u;
			const bool		e_ok = hash_has(a.hash, a.hash_cap, support_key(a.gen, r, q[0], es))
				|| hash_has(a.hash, a.hash_cap, support_key(a.gen, r, q[1], es));
			const bool f_ok = hash_has(a.hash, a.hash_cap, support_key(a.gen, r, q[2], fs))
				|| hash_has(a.hash, a.hash_cap, support_key(a.gen, r, q[3], fs))
				|| hash_has(a.hash, a.hash_cap, support_key(a.gen, r, q[4], fs));
			cnt += (e_ok && f_ok) ? 1u : 0u;
		}
		__shared__ uint32_t part[16];
		for (int o = 32; o > 0; o >>= 1)
			cnt += __shfl_xor(cnt, o, kWave);
		if (__lane_id() == 0)
			part[threadIdx.x / kWave] = cnt;
		__syncthreads();
		if (threadIdx.x == 0)
		{
			uint32_t t = 0;
			for (uint32_t k = 0; k < blockDim.x / kWave; ++k)
				t += part[k];
			a.cov[r] = t;
			atomicMax((unsigned long long*)&a.slot_best[2 * h.slot], ((unsigned long long)t << 32) | a.count[r]);
		}
		__syncthreads();
	}
}

// best cluster per slot: max coverage, then max hits, then min canonical key
__global__ void k_prp_select_key(CompArgs a)
{
	const uint32_t nr = *a.nroots;
	const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nr)
		return;
	const uint32_t r = a.roots[b];
	const uint32_t slot = a.h[r].slot;
	if ((((uint64_t)a.cov[r] << 32) | a.count[r]) == a.slot_best[2 * slot])
		atomicMin((unsigned long long*)&a.slot_best[2 * slot + 1], (unsigned long long)a.key[r]);
}

__global__ void k_prp_init_keys(uint32_t ns, uint64_t* best)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < ns)
		best[2 * j + 1] = ~0ull;
}

__global__ void k_prp_select_root(CompArgs a)
{
	const uint32_t nr = *a.nroots;
	const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nr)
		return;
	const uint32_t r = a.roots[b];
	const uint32_t slot = a.h[r].slot;
	if ((((uint64_t)a.cov[r] << 32) | a.count[r]) == a.slot_best[2 * slot] && a.key[r] == a.slot_best[2 * slot + 1])
		a.slot_root[slot] = r;
}

// ------------------------------------------------------------------ J(r) per slot (:1883-1954, :1806-1881)
struct EvalArgs
{
	uint32_t		nslots, V, c0;
	const PrpHit*	h;
	const uint32_t* range; // [2 * nslots]
	const uint32_t* slot_root;
	const uint32_t* cov;
	const PrpSlot*	slots;
	const uint32_t *cpair_off, *edges, *faces, *hop_all;
	const float *	va, *ea, *fa, *mass, *fn, *en, *vn;
	PrpSummary*		out;
};

__device__ __forceinline__ float inv_mass_of(const EvalArgs& a, uint32_t v) { return 1.0f / fmaxf(a.mass[v], 1e-8f); }

// Partial sums of one slot over one of kEvalSplit interleaved shares of its
// hits (so a slot with many hits is spread over several workgroups).
struct EvalPart
{
	double	 area, vol, cost;
	float	 depth;
	uint32_t cnt, hs, hd, pad;
};
constexpr uint32_t kEvalSplit = 16;

__global__ __launch_bounds__(256) void k_prp_eval(EvalArgs a, EvalPart* __restrict__ part)
{
	const uint32_t j = blockIdx.x, b = blockIdx.y;
	if (j >= a.nslots)
		return;
	const uint32_t lo = a.range[2 * j], hi = a.range[2 * j + 1];
	const uint32_t root = a.slot_root[j];
	const PrpSlot  sl = a.slots[j];
	const f3	   dir = mk3(sl.dir[0], sl.dir[1], sl.dir[2]);
	const uint32_t* H0 = a.hop_all + (size_t)(2 * sl.contour) * a.V;
	const uint32_t* H1 = a.hop_all + (size_t)(2 * sl.contour + 1) * a.V;
	// per-hit terms in FP32, sums in FP64 in a fixed order, rounded once
	double		   area = 0.0, vol = 0.0, cost = 0.0;
	float		   depth = 0.0f;
	uint32_t	   cnt = 0, hs = 0, hd = 0;
	if (root != kNone)
		for (uint32_t i = lo + b * blockDim.x + threadIdx.x; i < hi; i += kEvalSplit * blockDim.x)
		{
			const PrpHit h = a.h[i];
			if (h.root != root)
				continue;
			++cnt;
			float ar;
			f3	  nn;
			if (h.type == 0)
				ar = a.va[h.src], nn = ld3(a.vn, h.src);
			else if (h.type == 1)
				ar = a.ea[h.src], nn = ld3(a.en, h.src);
			else
				ar = a.fa[h.src], nn = ld3(a.fn, h.src);
			area += (double)ar;
			vol += (double)(ar * fabsf(dot3(nn, dir)) * h.dist);
			depth = fmaxf(depth, h.dist);
			float im;
			if (h.type == 1)
			{
				const float x0 = h.b0, x1 = 1.0f - h.b0, y0 = h.b1, y1 = 1.0f - h.b1;
				im = x0 * x0 * inv_mass_of(a, a.edges[2 * h.src]) + x1 * x1 * inv_mass_of(a, a.edges[2 * h.src + 1])
					+ y0 * y0 * inv_mass_of(a, a.edges[2 * h.dst]) + y1 * y1 * inv_mass_of(a, a.edges[2 * h.dst + 1]);
			}
			else
			{
				const uint32_t f = h.type == 0 ? h.dst : h.src, v = h.type == 0 ? h.src : h.dst;
				const float	   bz = 1.0f - h.b0 - h.b1;
				const float	   fpart = h.b0 * h.b0 * inv_mass_of(a, a.faces[3 * f])
					+ h.b1 * h.b1 * inv_mass_of(a, a.faces[3 * f + 1]) + bz * bz * inv_mass_of(a, a.faces[3 * f + 2]);
				im = h.type == 0 ? inv_mass_of(a, v) + fpart : fpart + inv_mass_of(a, v);
			}
			const float shd = h.dist + 3e-3f;
			cost += (double)(im > 1e-8f ? shd * shd / im : 0.0f);
			uint32_t	   sv[3], dv[3];
			const uint32_t ns = src_verts(a.edges, a.faces, h, sv), nd = dst_verts(a.edges, a.faces, h, dv);
			for (uint32_t k = 0; k < ns; ++k)
				if (H0[sv[k]] != kHopInf)
					hs = max(hs, H0[sv[k]]);
			for (uint32_t k = 0; k < nd; ++k)
				if (H1[dv[k]] != kHopInf)
					hd = max(hd, H1[dv[k]]);
		}
	area = wave_sum(area);
	vol = wave_sum(vol);
	cost = wave_sum(cost);
	depth = wave_max(depth);
	for (int o = 32; o > 0; o >>= 1)
	{
		cnt += __shfl_xor(cnt, o, kWave);
		hs = max(hs, (uint32_t)__shfl_xor(hs, o, kWave));
		hd = max(hd, (uint32_t)__shfl_xor(hd, o, kWave));
	}
	__shared__ double	sf[3][4];
	__shared__ float	sdepth[4];
	__shared__ uint32_t su[3][4];
	const uint32_t		w = threadIdx.x / kWave;
	if (__lane_id() == 0)
		sf[0][w] = area, sf[1][w] = vol, sf[2][w] = cost, sdepth[w] = depth, su[0][w] = cnt, su[1][w] = hs,
		su[2][w] = hd;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		EvalPart p{};
		for (uint32_t k = 0; k < blockDim.x / kWave; ++k)
		{
			p.area += sf[0][k], p.vol += sf[1][k], p.cost += sf[2][k];
			p.depth = fmaxf(p.depth, sdepth[k]);
			p.cnt += su[0][k];
			p.hs = max(p.hs, su[1][k]);
			p.hd = max(p.hd, su[2][k]);
		}
		part[(size_t)j * kEvalSplit + b] = p;
	}
}

__global__ void k_prp_eval_final(EvalArgs a, const EvalPart* __restrict__ part)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= a.nslots)
		return;
	const uint32_t lo = a.range[2 * j], hi = a.range[2 * j + 1];
	const uint32_t root = a.slot_root[j];
	const PrpSlot  sl = a.slots[j];
	PrpSummary	   s{};
	double		   sa = 0.0, sv = 0.0, sc = 0.0;
	for (uint32_t b = 0; b < kEvalSplit; ++b)
	{
		const EvalPart& p = part[(size_t)j * kEvalSplit + b];
		sa += p.area, sv += p.vol, sc += p.cost;
		s.depth = fmaxf(s.depth, p.depth);
		s.hit_count += p.cnt;
		s.max_hop_src = max(s.max_hop_src, p.hs);
		s.max_hop_dst = max(s.max_hop_dst, p.hd);
	}
	s.area = (float)sa, s.volume = (float)sv, s.disp_cost = (float)sc;
	s.raw_hits = hi - lo;
	s.ef_count = a.cpair_off[a.c0 + sl.contour + 1] - a.cpair_off[a.c0 + sl.contour];
	s.covered = root != kNone ? a.cov[root] : 0u;
	s.objective = s.area * s.depth * s.depth;
	s.root = root;
	const float cv = s.ef_count ? (float)s.covered / (float)s.ef_count : 0.0f;
	s.usable = s.hit_count > 0 && s.covered > 0 && isfinite(s.objective) && isfinite(cv) && cv > 0.0f;
	a.out[j] = s;
}

// two-pass LSD order (slot, type, src, dst): first by dst, then stably by the
// bucket key.  The raw hits sit in kHitLists sub-lists; sub-list k's entries
// take positions off_k .. off_k + count_k - 1 of the sort, and the values are
// the hits' indices in the sub-list storage.
__global__ void k_prp_dst_keys(const uint32_t* __restrict__ list_count, uint32_t sub_cap,
	const PrpHit* __restrict__ h, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
	const uint32_t k = blockIdx.y;
	__shared__ uint32_t s_off;
	if (threadIdx.x == 0)
	{
		uint32_t o = 0;
		for (uint32_t j = 0; j < k; ++j)
			o += list_count[kListStride * j];
		s_off = o;
	}
	__syncthreads();
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= list_count[kListStride * k])
		return;
	const uint32_t src = k * sub_cap + i;
	keys[s_off + i] = h[src].dst;
	vals[s_off + i] = src;
}

__global__ void k_prp_keys(uint32_t n, const PrpHit* __restrict__ h, const uint32_t* __restrict__ order, uint32_t pb,
	uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint32_t j = order[i];
	keys[i] = hit_key(h[j].slot, h[j].type, h[j].src, pb);
	vals[i] = j;
}

__global__ void k_prp_gather(uint32_t n, const PrpHit* __restrict__ in, const uint32_t* __restrict__ perm,
	PrpHit* __restrict__ out, uint32_t* __restrict__ dst, uint32_t nslots, uint32_t* __restrict__ range)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const PrpHit h = in[perm[i]];
	out[i] = h;
	dst[i] = h.dst;
	const uint32_t prev = i > 0 ? in[perm[i - 1]].slot : kNone;
	const uint32_t next = i + 1 < n ? in[perm[i + 1]].slot : kNone;
	if (prev != h.slot)
		range[2 * h.slot] = i;
	if (next != h.slot)
		range[2 * h.slot + 1] = i + 1;
}

__global__ void k_prp_fill_u32(uint32_t n, uint32_t* p, uint32_t v)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}
__global__ void k_prp_fill_u64(uint32_t n, uint64_t* p, uint64_t v)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}
__global__ void k_prp_fill_u8(uint32_t n, uint8_t* p, uint8_t v)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

// ------------------------------------------------------------------ contour adjacency + responses
// get_adj_contours (:414-448): another contour o whose flags cover every source vertex
// with 2o + s and every target vertex with 2o + s ^ 1 (s in {0, 1})
__global__ void k_prp_adjacent_contours(uint32_t n, const PrpHit* __restrict__ h, const uint32_t* __restrict__ hc,
	const uint32_t* __restrict__ flag_off, const uint32_t* __restrict__ flags, const uint32_t* __restrict__ edges,
	const uint32_t* __restrict__ faces, uint32_t nc, uint8_t* __restrict__ adjm)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const PrpHit   x = h[i];
	const uint32_t c = hc[i];
	uint32_t	   sv[3], dv[3];
	const uint32_t ns = src_verts(edges, faces, x, sv), nd = dst_verts(edges, faces, x, dv);
	for (uint32_t k = flag_off[sv[0]]; k < flag_off[sv[0] + 1]; ++k)
	{
		const uint32_t fs = flags[k];
		if (fs / 2 == c)
			continue;
		bool ok = true;
		for (uint32_t q = 1; q < ns && ok; ++q)
			ok = has_flag(flag_off, flags, sv[q], fs);
		for (uint32_t q = 0; q < nd && ok; ++q)
			ok = has_flag(flag_off, flags, dv[q], fs ^ 1u);
		if (ok)
			adjm[(size_t)c * nc + fs / 2] = 1u;
	}
}

// make_response_from_min_hitinfo (:2105-2221): one CollisionPairTemplate per kept hit
// of an unsuppressed contour, at its rank among those (flags + exclusive scan: the
// records keep the final list's order)
__global__ void k_prp_response_flags(uint32_t n, const uint32_t* __restrict__ hc,
	const uint8_t* __restrict__ suppressed, uint32_t* __restrict__ flag)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		flag[i] = suppressed[hc[i]] ? 0u : 1u;
}

__global__ void k_prp_responses(uint32_t n, const PrpHit* __restrict__ h, const uint32_t* __restrict__ hc,
	const float* __restrict__ cdir, const float* __restrict__ cC, const uint8_t* __restrict__ suppressed,
	const uint32_t* __restrict__ edges, const uint32_t* __restrict__ faces, const float* __restrict__ va,
	const float* __restrict__ ea, const float* __restrict__ fa, float kappa, const uint32_t* __restrict__ pos,
	ContactRec* __restrict__ out, uint4* __restrict__ out_meta, uint32_t cap)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || suppressed[hc[i]])
		return;
	const uint32_t slot = pos[i];
	if (slot >= cap)
		return;
	const PrpHit   x = h[i];
	const uint32_t c = hc[i];
	ContactRec	   r{};
	float		   nn[3] = { cdir[3 * c], cdir[3 * c + 1], cdir[3 * c + 2] };
	float		   ar;
	if (x.type == 0)
	{
		ar = fmaxf(0.5f * (va[x.src] + fa[x.dst]), 3e-6f);
		r.idx[0] = x.src, r.idx[1] = faces[3 * x.dst], r.idx[2] = faces[3 * x.dst + 1] | 0x80000000u;
		r.idx[3] = faces[3 * x.dst + 2];
	}
	else if (x.type == 2)
	{
		ar = fmaxf(0.5f * (fa[x.src] + va[x.dst]), 3e-6f);
		r.idx[0] = x.dst, r.idx[1] = faces[3 * x.src], r.idx[2] = faces[3 * x.src + 1] | 0x80000000u;
		r.idx[3] = faces[3 * x.src + 2];
		nn[0] = -nn[0], nn[1] = -nn[1], nn[2] = -nn[2];
	}
	else
	{
		ar = fmaxf(0.5f * (ea[x.src] + ea[x.dst]), 3e-6f);
		r.idx[0] = edges[2 * x.src], r.idx[1] = edges[2 * x.src + 1];
		r.idx[2] = edges[2 * x.dst], r.idx[3] = edges[2 * x.dst + 1] | 0x80000000u; // make_ee_pair type bit
	}
	const float stiff = ar * kappa;
	r.normal[0] = nn[0], r.normal[1] = nn[1], r.normal[2] = nn[2];
	r.area = ar;
	r.k1 = stiff * cC[c];
	r.k2 = stiff;
	r.bary[0] = x.b0, r.bary[1] = x.b1;
	out[slot] = r;
	// target_point_template_pairs_indices (:2159, :2178, :2196): VF (vertex, face, region, 0),
	// FV (vertex, face, region, 1), EE (source edge, target edge, region, 2); region =
	// prp_pack_response_region_label(c, 0, label) = 2 * 256 * c with display label 0
	const uint32_t region = 2u * 256u * c;
	out_meta[slot] = x.type == 0 ? make_uint4(x.src, x.dst, region, 0u)
		: x.type == 2			 ? make_uint4(x.dst, x.src, region, 1u)
								 : make_uint4(x.src, x.dst, region, 2u);
}

// The kept hits of the chosen slots (root == slot root) go to the final list in
// their sorted order: flags, an exclusive scan, then the copy (no atomics, so the
// final list -- and the response records made from it -- has one fixed order).
__global__ void k_prp_keep_flags(uint32_t n, const PrpHit* __restrict__ h, const uint32_t* __restrict__ slot_root,
	uint32_t* __restrict__ flag)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		flag[i] = h[i].root == slot_root[h[i].slot] ? 1u : 0u;
}

__global__ void k_prp_keep(uint32_t n, const PrpHit* __restrict__ h, const uint32_t* __restrict__ slot_root,
	const PrpSlot* __restrict__ slots, uint32_t c0, const uint32_t* __restrict__ pos, uint32_t base,
	PrpHit* __restrict__ out, uint32_t* __restrict__ out_c, uint32_t cap)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && h[i].root == slot_root[h[i].slot] && base + pos[i] < cap)
	{
		out[base + pos[i]] = h[i];
		out_c[base + pos[i]] = c0 + slots[h[i].slot].contour;
	}
}

} // namespace

// ================================================================== host side

namespace
{

void build_csr(uint32_t n, const std::vector<uint32_t>& src, const std::vector<uint32_t>& dst,
	std::vector<uint32_t>& off, std::vector<uint32_t>& val, bool unique)
{
	off.assign(n + 1, 0);
	for (uint32_t s : src)
		off[s + 1]++;
	for (uint32_t i = 0; i < n; ++i)
		off[i + 1] += off[i];
	val.resize(src.size());
	std::vector<uint32_t> fill(off.begin(), off.end() - 1);
	for (size_t i = 0; i < src.size(); ++i)
		val[fill[src[i]]++] = dst[i];
	uint32_t w = 0;
	for (uint32_t i = 0; i < n; ++i)
	{
		const uint32_t b = off[i], e = off[i + 1];
		std::sort(val.begin() + b, val.begin() + e);
		const uint32_t row = w;
		for (uint32_t k = b; k < e; ++k)
			if (!unique || w == row || val[w - 1] != val[k])
				val[w++] = val[k];
		off[i] = row;
	}
	off[n] = w;
	val.resize(w);
}

// Synchronous upload: the host vector may be reused or freed right after the
// call, so the copy must have consumed it before returning.
template <typename T>
void upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s)
{
	d.reserve(h.size());
	if (!h.empty())
	{
		UT_HIP_CHECK(hipMemcpyAsync(d.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
	}
}

// symmetric 3x3 eigenvectors (cyclic Jacobi in FP64), ascending eigenvalues,
// sign fixed so the largest-magnitude component is positive
void eigen_sym3(const double S[9], double vec[3][3])
{
	double A[9], V[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
	std::memcpy(A, S, sizeof(A));
	for (int sweep = 0; sweep < 64; ++sweep)
	{
		if (A[1] * A[1] + A[2] * A[2] + A[5] * A[5] < 1e-300)
			break;
		for (int p = 0; p < 2; ++p)
			for (int q = p + 1; q < 3; ++q)
			{
				const double apq = A[3 * p + q];
				if (std::fabs(apq) < 1e-300)
					continue;
				const double th = (A[3 * q + q] - A[3 * p + p]) / (2.0 * apq);
				const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
				const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
				for (int k = 0; k < 3; ++k)
				{
					const double kp = A[3 * k + p], kq = A[3 * k + q];
					A[3 * k + p] = c * kp - s * kq;
					A[3 * k + q] = s * kp + c * kq;
				}
				for (int k = 0; k < 3; ++k)
				{
					const double pk = A[3 * p + k], qk = A[3 * q + k];
					A[3 * p + k] = c * pk - s * qk;
					A[3 * q + k] = s * pk + c * qk;
				}
				for (int k = 0; k < 3; ++k)
				{
					const double kp = V[3 * k + p], kq = V[3 * k + q];
					V[3 * k + p] = c * kp - s * kq;
					V[3 * k + q] = s * kp + c * kq;
				}
			}
	}
	int idx[3] = { 0, 1, 2 };
	std::sort(idx, idx + 3, [&](int a, int b) { return A[4 * a] < A[4 * b]; });
	for (int i = 0; i < 3; ++i)
	{
		double v[3] = { V[idx[i]], V[3 + idx[i]], V[6 + idx[i]] };
		int	   big = 0;
		for (int k = 1; k < 3; ++k)
			if (std::fabs(v[k]) > std::fabs(v[big]))
				big = k;
		const double sg = v[big] < 0 ? -1.0 : 1.0;
		for (int k = 0; k < 3; ++k)
			vec[i][k] = sg * v[k];
	}
}

void push_dir(std::vector<float>& out, const double* d, double sign)
{
	if (!(std::isfinite(d[0]) && std::isfinite(d[1]) && std::isfinite(d[2])))
		return;
	const double l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
	if (!(l2 > 1e-12) || out.size() >= 3 * kMaxDirs)
		return;
	const double l = std::sqrt(l2);
	for (int k = 0; k < 3; ++k)
		out.push_back((float)(sign * d[k] / l));
}

uint32_t initial_k(uint32_t nb, uint32_t diameter)
{
	if (diameter == 0)
		return 3;
	const uint32_t ka = (uint32_t)std::ceil(std::sqrt((float)nb));
	return std::max(3u, std::min(ka, diameter));
}

uint32_t grow_k(uint32_t cur, uint32_t diam)
{
	if (cur >= diam)
		return cur;
	return (uint32_t)std::min<uint64_t>(diam, std::max<uint64_t>(1u, (uint64_t)cur * 2u));
}

struct ComboState
{
	bool	   present = false, converged = false, pruned = false, eligible = false;
	PrpSummary s{};
	uint32_t   ks = 0, kd = 0;
};

float response_objective(const PrpSummary& s)
{
	const float cov = s.ef_count ? (float)s.covered / (float)s.ef_count : 0.0f;
	if (!std::isfinite(s.objective) || !std::isfinite(cov) || cov <= 0.0f)
		return INFINITY;
	const double o = (double)s.objective / ((double)cov * cov);
	return std::isfinite(o) && o <= (double)FLT_MAX ? (float)o : INFINITY;
}

int scalar_cmp(float l, float r)
{
	const bool lf = std::isfinite(l), rf = std::isfinite(r);
	if (lf != rf)
		return lf ? -1 : 1;
	if (!lf)
		return 0;
	const float tol = 1e-7f + 1e-5f * std::max(std::fabs(l), std::fabs(r));
	if (l + tol < r)
		return -1;
	if (r + tol < l)
		return 1;
	return 0;
}

// prp_response_is_better (intersection_resolver2.h:852-884)
bool better(const PrpSummary& a, uint32_t ca, const PrpSummary& b, uint32_t cb)
{
	int c = scalar_cmp(response_objective(a), response_objective(b));
	if (c)
		return c < 0;
	if ((c = scalar_cmp(a.depth, b.depth)))
		return c < 0;
	if ((c = scalar_cmp(a.disp_cost, b.disp_cost)))
		return c < 0;
	if (a.hit_count != b.hit_count)
		return a.hit_count < b.hit_count;
	return ca < cb;
}

} // namespace

void PrpResolver::init(Context* c, uint32_t nv_, uint32_t ne_, uint32_t nf_, const uint32_t* edges_,
	const uint32_t* faces_, const uint32_t* fae_, const uint32_t* eaf_, const float* rest_x_, const float* va_,
	const float* ea_, const float* fa_, const float* mass_, const PrpParams& p)
{
	ctx = c;
	nv = nv_, ne = ne_, nf = nf_;
	edges = edges_, faces = faces_, fae = fae_, eaf = eaf_, rest_x = rest_x_;
	va = va_, ea = ea_, fa = fa_, mass = mass_;
	params = p;
	UT_REQUIRE(p.total_dirs >= 1 && p.total_dirs <= kMaxDirs, "prp: total_dirs must be in [1, 64]");
	UT_REQUIRE(p.batch_contours >= 1, "prp: batch_contours must be positive");
	const uint32_t mx = std::max({ nv, ne, nf, 2u });
	prim_bits = 32 - __builtin_clz(mx - 1);
	UT_REQUIRE(prim_bits <= 26, "prp: mesh too large for 32-bit hit keys");
	slot_bits = 30 - prim_bits;
	const hipStream_t	  s = ctx->stream;
	std::vector<uint32_t> he(2 * (size_t)ne), hf(3 * (size_t)nf);
	UT_HIP_CHECK(hipMemcpyAsync(he.data(), edges, 8 * (size_t)ne, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipMemcpyAsync(hf.data(), faces, 12 * (size_t)nf, hipMemcpyDeviceToHost, s));
	h_va.resize(nv);
	UT_HIP_CHECK(hipMemcpyAsync(h_va.data(), va, 4 * (size_t)nv, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	std::vector<uint32_t> src, dst, off, val;
	for (uint32_t e = 0; e < ne; ++e)
		src.push_back(he[2 * e]), dst.push_back(he[2 * e + 1]), src.push_back(he[2 * e + 1]), dst.push_back(he[2 * e]);
	build_csr(nv, src, dst, off, val, true);
	upload(vv_off, off, s), upload(vv, val, s);
	src.clear(), dst.clear();
	for (uint32_t e = 0; e < ne; ++e)
		src.push_back(he[2 * e]), dst.push_back(e), src.push_back(he[2 * e + 1]), dst.push_back(e);
	build_csr(nv, src, dst, off, val, true);
	upload(ve_off, off, s), upload(ve, val, s);
	src.clear(), dst.clear();
	for (uint32_t f = 0; f < nf; ++f)
		for (int k = 0; k < 3; ++k)
			src.push_back(hf[3 * f + k]), dst.push_back(f);
	build_csr(nv, src, dst, off, val, true);
	upload(vf_off, off, s), upload(vf, val, s);
	normals.reserve(3 * ((size_t)nf + ne + nv));
	counters.reserve(8);
	status.reserve(2);
	spill.reserve((size_t)kRayBlocks * kTravThreads * (kStackTotal - kRayStackLds));
	tstat.reserve(16);
	if (!ee_ev[0])
		for (auto& e : ee_ev)
			UT_HIP_CHECK(hipEventCreate(&e));
	UT_HIP_CHECK(hipStreamSynchronize(s));
}

PrpResolver::~PrpResolver()
{
	for (auto& e : ee_ev)
		if (e)
			(void)hipEventDestroy(e);
}

void PrpResolver::build_contour_tables(const ContourBuilder& ct)
{
	const hipStream_t s = ctx->stream;
	total_pairs = ct.num_total;
	nc = ct.num_contours;
	const uint32_t		  N = total_pairs;
	std::vector<uint32_t> words(12 * (size_t)N), cidx(N), midx(N), adjn(N), adj(ct.num_adj), adj_off(N + 1);
	if (N)
	{
		UT_HIP_CHECK(hipMemcpyAsync(words.data(), ct.pairs.get(), 48 * (size_t)N, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipMemcpyAsync(cidx.data(), ct.contour_index.get(), 4 * (size_t)N, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipMemcpyAsync(midx.data(), ct.mesh_index.get(), 4 * (size_t)N, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipMemcpyAsync(adjn.data(), ct.adj_n.get(), 4 * (size_t)N, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipMemcpyAsync(adj_off.data(), ct.adj_off.get(), 4 * (size_t)(N + 1), hipMemcpyDeviceToHost, s));
		if (ct.num_adj)
			UT_HIP_CHECK(hipMemcpyAsync(adj.data(), ct.adj.get(), 4 * (size_t)ct.num_adj, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
	}
	h_pair_words = words;
	h_contour_of_pair = cidx;
	h_mesh_of_pair = midx;
	h_adj = adj;
	h_adjn = adjn;
	h_adj_off = adj_off;
	// pairs of each contour, ascending pair index
	h_cpair_off.assign(nc + 1, 0);
	for (uint32_t p = 0; p < N; ++p)
		h_cpair_off[cidx[p] + 1]++;
	for (uint32_t c = 0; c < nc; ++c)
		h_cpair_off[c + 1] += h_cpair_off[c];
	h_cpairs.resize(N);
	std::vector<uint32_t> fill(h_cpair_off.begin(), h_cpair_off.end() - 1);
	for (uint32_t p = 0; p < N; ++p)
		h_cpairs[fill[cidx[p]]++] = p;
	std::vector<uint32_t> rec(6 * (size_t)N);
	for (uint32_t i = 0; i < N; ++i)
	{
		const uint32_t p = h_cpairs[i];
		for (int k = 0; k < 5; ++k)
			rec[6 * (size_t)i + k] = words[12 * (size_t)p + k];
		rec[6 * (size_t)i + 5] = midx[p];
	}
	// boundary vertices per (contour, role) from every pair; flags from non-loop pairs
	std::vector<uint32_t> bs, bd, fs, fd;
	for (uint32_t p = 0; p < N; ++p)
	{
		const uint32_t* q = &words[12 * (size_t)p];
		const uint32_t	m = midx[p], c = cidx[p];
		for (int k = 0; k < 5; ++k)
			bs.push_back(2 * c + (k < 2 ? m : m ^ 1u)), bd.push_back(q[k]);
		if (q[0] == q[1])
			continue;
		for (int k = 0; k < 5; ++k)
			fs.push_back(q[k]), fd.push_back(2 * c + (k < 2 ? m : m ^ 1u));
	}
	std::vector<uint32_t> foff, fval;
	build_csr(2 * nc, bs, bd, h_b_off, h_b_verts, true);
	build_csr(nv, fs, fd, foff, fval, true);
	upload(cpair_off, h_cpair_off, s);
	upload(cpair_rec, rec, s);
	upload(b_off, h_b_off, s);
	upload(b_verts, h_b_verts, s);
	upload(flag_off, foff, s);
	upload(flags, fval, s);
	if (fval.empty())
	{
		flags.reserve(1);
	}
}

void PrpResolver::directions(uint32_t c0, uint32_t n)
{
	// global area-weighted centroid (computed once per resolve)
	const double* g = global_c;
	for (uint32_t lc = 0; lc < n; ++lc)
	{
		const uint32_t c = c0 + lc;
		const uint32_t p0 = h_cpair_off[c], p1 = h_cpair_off[c + 1], np = p1 - p0;
		auto		   X = [&](uint32_t v, int k) { return h_x[3 * (size_t)v + k]; };
		auto		   pos = [&](uint32_t p, float out[3])
		{
			const uint32_t* q = &h_pair_words[12 * (size_t)p];
			float			b0;
			std::memcpy(&b0, q + 5, 4);
			for (int k = 0; k < 3; ++k)
				out[k] = X(q[0], k) * b0 + X(q[1], k) * (1.0f - b0);
		};
		double c_[3] = { 0, 0, 0 }, u[3] = { 0, 0, 0 }, sw = 0;
		for (uint32_t i = p0; i < p1; ++i)
		{
			const uint32_t p = h_cpairs[i];
			float		   pp[3], w = 0.0f;
			pos(p, pp);
			for (uint32_t j = 0; j < h_adjn[p]; ++j)
			{
				float qq[3];
				pos(h_adj[h_adj_off[p] + j], qq);
				const float dx = qq[0] - pp[0], dy = qq[1] - pp[1], dz = qq[2] - pp[2];
				w += std::sqrt(dx * dx + dy * dy + dz * dz);
			}
			w *= 0.5f;
			for (int k = 0; k < 3; ++k)
				c_[k] += (double)pp[k] * w, u[k] += pp[k];
			sw += w;
		}
		for (int k = 0; k < 3; ++k)
			c_[k] = sw == 0.0 ? u[k] / np : c_[k] / sw;
		double S[9] = { 0 }, mc[2][3] = { { 0 } };
		for (int r = 0; r < 2; ++r)
		{
			double ws = 0;
			for (uint32_t i = h_b_off[2 * c + r]; i < h_b_off[2 * c + r + 1]; ++i)
			{
				const uint32_t v = h_b_verts[i];
				const double   w = std::max((double)h_va[v], 1e-8);
				double		   d[3];
				for (int k = 0; k < 3; ++k)
					mc[r][k] += (double)X(v, k) * w, d[k] = X(v, k) - c_[k];
				for (int a = 0; a < 3; ++a)
					for (int b = 0; b < 3; ++b)
						S[3 * a + b] += w * d[a] * d[b];
				ws += w;
			}
			for (int k = 0; k < 3; ++k)
				mc[r][k] = ws > 1e-8 ? mc[r][k] / ws : c_[k];
		}
		double E[3][3] = { { 0, 1, 0 }, { 1, 0, 0 }, { 0, 0, 1 } };
		double tr = S[0] + S[4] + S[8], fro = 0;
		bool   ok = true;
		for (int k = 0; k < 9; ++k)
			fro += S[k] * S[k], ok = ok && std::isfinite(S[k]);
		if (ok && tr >= 1e-10 && fro >= 1e-20)
			eigen_sym3(S, E);
		std::vector<float> d;
		for (int i = 0; i < 3; ++i)
			push_dir(d, E[i], 1.0), push_dir(d, E[i], -1.0);
		for (int r = 0; r < 2; ++r)
		{
			double dd[3] = { c_[0] - mc[r][0], c_[1] - mc[r][1], c_[2] - mc[r][2] };
			const double l2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
			if (!(std::isfinite(l2) && l2 > 1e-12))
				std::memcpy(dd, E[r + 1], sizeof(dd));
			push_dir(d, dd, 1.0);
		}
		const double dg[3] = { c_[0] - g[0], c_[1] - g[1], c_[2] - g[2] };
		push_dir(d, dg, 1.0), push_dir(d, dg, -1.0);
		const uint32_t total = params.total_dirs;
		for (uint32_t i = 0; total > 10 && d.size() / 3 < total && i < 4 * kMaxDirs; ++i)
		{
			const uint32_t m = total > 10 ? total - 10 : 1;
			const double   zz = 1.0 - (2.0 * i + 1.0) / m;
			const double   rr = std::sqrt(std::max(0.0, 1.0 - zz * zz));
			const double   ph = i * 2.399963229728653;
			const double   f[3] = { rr * std::cos(ph), zz, rr * std::sin(ph) };
			push_dir(d, f, 1.0);
		}
		h_dirs[c] = d;
	}
}

// Distance state for the (contour, role) slots of n contours: rows, label words,
// work lists, thresholds, and the neighbour blocks (rest lengths).
void PrpResolver::dist_reserve(uint32_t n)
{
	const hipStream_t s = ctx->stream;
	const size_t	  V = nv;
	const size_t	  nn = std::max<size_t>(n, params.batch_contours);
	hop.reserve(2 * nn * V);
	geo.reserve(2 * nn * V);
	queue.reserve(8 * nn * V);
	inq.reserve(2 * nn * V);
	own.reserve(2 * (size_t)params.batch_contours * V);
	dist_state.reserve(4 * nn);
	dist_out.reserve(4 * nn);
	adj_blocks.reserve((size_t)2 * kAdjWays * V);
	k_prp_adj_blocks<<<div_up(nv, 256), 256, 0, s>>>(nv, vv_off.get(), vv.get(), rest_x,
		reinterpret_cast<uint2*>(adj_blocks.get()));
	UT_LAUNCH_CHECK();
}

// One distance launch over the listed slots of contours from c0 (triples: slot,
// hop threshold, first launch); mode B raises the geodesic radii instead.
void PrpResolver::dist_pass(uint32_t c0, const std::vector<uint32_t>& req, bool mode_b)
{
	if (req.empty())
		return;
	const hipStream_t s = ctx->stream;
	upload(dist_req, req, s);
	const uint32_t bw = dist_bitmap_words(nv);
	const size_t   lds = 4 * (2 * (size_t)kDistListLds + 2 * (size_t)bw);
	DistArgs	   da{ nv, c0, dist_req.get(), mode_b ? 1u : 0u, b_off.get(), b_verts.get(), vv_off.get(), vv.get(), rest_x,
		  reinterpret_cast<const uint2*>(adj_blocks.get()), hop.get(), geo.get(), queue.get(), inq.get(),
		  reinterpret_cast<uint2*>(dist_state.get()), reinterpret_cast<uint2*>(dist_out.get()), bw };
	k_prp_distances<<<(uint32_t)(req.size() / 3), kBfsThreads, lds, s>>>(da);
	UT_LAUNCH_CHECK();
}

// (largest finite hop, ring geodesic bound) of the first 2n slots
void PrpResolver::dist_results(uint32_t n)
{
	const hipStream_t s = ctx->stream;
	h_dist_out.resize(4 * (size_t)n);
	UT_HIP_CHECK(hipMemcpyAsync(h_dist_out.data(), dist_out.get(), 16 * (size_t)n, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
}

// full hop and geodesic fields of contours [c0, c0 + n) in batch rows 0 .. 2n - 1
void PrpResolver::distances(uint32_t c0, uint32_t n)
{
	dist_reserve(n);
	std::vector<uint32_t> req;
	for (uint32_t sl = 0; sl < 2 * n; ++sl)
		req.insert(req.end(), { sl, kDistHopAll, 1u });
	dist_pass(c0, req, false);
	dist_results(n);
	h_diam.resize(2 * n);
	for (uint32_t sl = 0; sl < 2 * n; ++sl)
		h_diam[sl] = h_dist_out[2 * sl];
}

void PrpResolver::prepare(const float* x, const ContourBuilder& ct, uint32_t c0, uint32_t n)
{
	const hipStream_t s = ctx->stream;
	x_dev = x;
	build_contour_tables(ct);
	UT_REQUIRE(c0 + n <= nc, "prp: contour range out of bounds");
	UT_REQUIRE(n <= params.batch_contours, "prp: batch larger than batch_contours");
	h_x.resize(3 * (size_t)nv);
	UT_HIP_CHECK(hipMemcpyAsync(h_x.data(), x, 12 * (size_t)nv, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	double gw = 0;
	global_c[0] = global_c[1] = global_c[2] = 0.0;
	for (uint32_t v = 0; v < nv; ++v)
	{
		const double w = std::max((double)h_va[v], 1e-8);
		for (int k = 0; k < 3; ++k)
			global_c[k] += w * h_x[3 * (size_t)v + k];
		gw += w;
	}
	for (int k = 0; k < 3; ++k)
		global_c[k] /= gw;
	float* fn = normals.get();
	float* en = fn + 3 * (size_t)nf;
	float* vn = en + 3 * (size_t)ne;
	k_prp_face_normals<<<div_up(nf, 256), 256, 0, s>>>(nf, x, faces, fn);
	k_prp_edge_vert_normals<<<div_up(ne + nv, 256), 256, 0, s>>>(ne, nv, eaf, vf_off.get(), vf.get(), fa, fn, en, vn);
	UT_LAUNCH_CHECK();
	h_dirs.assign(nc, {});
	batch0 = c0, batch_n = n;
	if (n == 0)
		return;
	distances(c0, n);
	begin_batch(0);
	directions(c0, n);
}

// The batch's (contour, role) rows start at contour `first` of the distance
// arrays; ownership is recomputed for the batch.
void PrpResolver::begin_batch(uint32_t first)
{
	const hipStream_t s = ctx->stream;
	hop_b = hop.get() + 2 * (size_t)first * nv;
	geo_b = geo.get() + 2 * (size_t)first * nv;
	diam_b = h_diam.data() + 2 * (size_t)first;
	const uint32_t m = 2 * batch_n * nv;
	k_prp_fill_u8<<<div_up(m, 256), 256, 0, s>>>(m, own.get(), (uint8_t)kOwnUnknown);
	UT_LAUNCH_CHECK();
}

void PrpResolver::ensure_hits(uint32_t cap)
{
	cap = div_up(cap, kHitLists) * kHitLists; // whole sub-lists
	if (cap <= hit_cap)
		return;
	hit_cap = cap;
	hits.reserve(cap), hits_sorted.reserve(cap), hit_dst.reserve(cap);
	for (int i = 0; i < 2; ++i)
		sort_k[i].reserve(cap), sort_v[i].reserve(cap);
	sort_h.reserve(radix_sort_hist_size(cap));
	comp_info.reserve(5 * (size_t)cap);
	comp_key.reserve(cap);
	roots.reserve(cap);
}

uint32_t PrpResolver::eval_slots(const std::vector<PrpSlot>& sl, Lbvh& face_tree, Lbvh& edge_tree,
	const std::vector<uint32_t>& kr, std::vector<PrpSummary>& out)
{
	const hipStream_t s = ctx->stream;
	const uint32_t	  ns = (uint32_t)sl.size(), n2 = 2 * batch_n, V = nv, E = ne, F = nf;
	out.assign(ns, PrpSummary{});
	if (ns == 0)
		return 0;
	UT_REQUIRE(ns <= (1u << slot_bits) && ns <= 4096, "prp: too many slots in one evaluation");
	UT_REQUIRE(kr.size() == n2, "prp: one k per (contour, role)");
	upload(slots, sl, s);
	upload(kring, kr, s);
	cand.reserve((size_t)n2 * (V + E));
	cand_cnt.reserve(2 * (size_t)n2);
	target_cnt.reserve(n2);
	req_off.reserve(3 * ((size_t)ns + 1));
	slot_range.reserve(2 * (size_t)ns);
	slot_best.reserve(2 * (size_t)ns);
	slot_root.reserve(ns);
	summary.reserve(ns);
	if (hit_cap == 0)
		ensure_hits(1u << 23); // 8M hits (256 MB): the 256^2 fold needs ~5M per evaluation
	// ownership and candidates
	k_prp_own<<<dim3(div_up(V, 256), n2), 256, 0, s>>>(
		V, batch0, n2, kring.get(), hop_b, geo_b, x_dev, b_off.get(), b_verts.get(), params.tau, own.get());
	k_prp_fill_u32<<<div_up(2 * n2, 256), 256, 0, s>>>(2 * n2, cand_cnt.get(), 0u);
	k_prp_fill_u32<<<div_up(n2, 256), 256, 0, s>>>(n2, target_cnt.get(), 0u);
	const uint32_t ewords = div_up(E, 64), fwords = div_up(F, 64);
	ecand.reserve((size_t)n2 * ewords);
	fcand.reserve((size_t)n2 * fwords);
	k_prp_cand_lists<<<dim3(div_up(std::max(V, E), 256), n2), 256, 0, s>>>(
		V, E, n2, kring.get(), hop_b, own.get(), edges, cand.get(), cand_cnt.get(), ecand.get(), ewords);
	k_prp_any_faces<<<dim3(div_up(F, 256), n2), 256, 0, s>>>(
		F, n2, kring.get(), V, hop_b, own.get(), faces, target_cnt.get(), fcand.get(), fwords);
	slot_tbox.reserve(12 * (size_t)ns);
	k_prp_slot_prep<<<ns, 256, 0, s>>>(V, E, ns, slots.get(), cand.get(), cand_cnt.get(), x_dev, edges,
		slot_tbox.get());
	k_prp_request_offsets<<<1, 1024, 0, s>>>(ns, slots.get(), cand_cnt.get(), target_cnt.get(), req_off.get());
	UT_LAUNCH_CHECK();
	uint32_t			  nh = 0, max_list = 0;
	std::vector<uint32_t> lc(kHitLists * kListStride);
	list_count.reserve(kHitLists * kListStride);
	for (int attempt = 0; attempt < 8; ++attempt)
	{
		RayArgs ra{ V, E, ns, slots.get(), req_off.get(), cand.get(), hop_b, own.get(), kring.get(), x_dev, edges,
			faces, hits.get(), list_count.get(), hit_cap / kHitLists, status.get(), spill.get(), nullptr, slot_tbox.get(),
			ecand.get(),
			fcand.get(), ewords, fwords, tstat.get() };
		k_prp_fill_u64<<<1, 64, 0, s>>>(16, reinterpret_cast<uint64_t*>(tstat.get()), 0ull);
		k_prp_fill_u32<<<div_up(kHitLists * kListStride, 256), 256, 0, s>>>(kHitLists * kListStride, list_count.get(), 0u);
		k_prp_fill_u32<<<1, 64, 0, s>>>(1, status.get(), 1u);	  // [0] tree health
		k_prp_fill_u32<<<1, 64, 0, s>>>(1, status.get() + 1, 0u); // [1] stack overflows
		ra.nodes = face_tree.nodes.get();
		k_prp_ray_vert<0><<<kRayBlocks, kTravThreads, 0, s>>>(ra);
		k_prp_ray_vert<2><<<kRayBlocks, kTravThreads, 0, s>>>(ra);
		ra.nodes = edge_tree.nodes.get();
		UT_HIP_CHECK(hipEventRecord(ee_ev[0], s));
		k_prp_ray_edge<<<kEeBlocks, kTravThreads, 0, s>>>(ra);
		UT_HIP_CHECK(hipEventRecord(ee_ev[1], s));
		UT_LAUNCH_CHECK();
		uint32_t		   st[2], tot[3];
		unsigned long long ts[16];
		UT_HIP_CHECK(hipMemcpyAsync(ts, tstat.get(), sizeof(ts), hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipMemcpyAsync(lc.data(), list_count.get(), 4 * lc.size(), hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipMemcpyAsync(st, status.get(), 8, hipMemcpyDeviceToHost, s));
		for (int t = 0; t < 3; ++t)
			UT_HIP_CHECK(hipMemcpyAsync(&tot[t], req_off.get() + (size_t)t * (ns + 1) + ns, 4, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		if (attempt == 0)
		{
			for (int t = 0; t < 3; ++t)
				ray_count[t] += tot[t];
			float ms = 0.0f;
			UT_HIP_CHECK(hipEventElapsedTime(&ms, ee_ev[0], ee_ev[1]));
			ee_prof.launches += 1;
			ee_prof.ms += ms;
			ee_prof.requests += tot[1];
			ee_prof.visits += ts[4];
			ee_prof.leaf_tests += ts[5]; // leaves reached by the sweep (each staged and tested)
			ee_prof.hits += ts[6];
			ee_prof.cand_tests += ts[7]; // of which candidates of role 1: the edge-edge test runs
			ee_prof.wave_steps += ts[13];
		}
		UT_REQUIRE(st[0] == 1u, "prp: ray traversal revisited a node (corrupt tree)");
		UT_REQUIRE(st[1] == 0u, "prp: ray traversal stack overflow");
		nh = 0, max_list = 0;
		for (uint32_t k = 0; k < kHitLists; ++k)
			nh += lc[kListStride * k], max_list = std::max(max_list, lc[kListStride * k]);
		if (max_list <= hit_cap / kHitLists)
			break;
		ensure_hits(std::max(2 * hit_cap, (max_list + max_list / 4) * kHitLists));
	}
	UT_REQUIRE(max_list <= hit_cap / kHitLists, "prp: hit buffer could not be grown");
	last_raw_hits = nh;
	k_prp_fill_u32<<<div_up(2 * ns, 256), 256, 0, s>>>(2 * ns, slot_range.get(), 0u);
	k_prp_fill_u32<<<div_up(ns, 256), 256, 0, s>>>(ns, slot_root.get(), kNone);
	k_prp_fill_u64<<<div_up(2 * ns, 256), 256, 0, s>>>(2 * ns, slot_best.get(), 0ull);
	UT_LAUNCH_CHECK();
	if (nh > 0)
	{
		// sort by (slot, type, source primitive, target primitive): by dst, then stably by the key
		k_prp_dst_keys<<<dim3(div_up(max_list, 256), kHitLists), 256, 0, s>>>(list_count.get(), hit_cap / kHitLists,
			hits.get(), sort_k[0].get(), sort_v[0].get());
		RadixSortBuffers r1;
		r1.keys[0] = sort_k[0].get(), r1.keys[1] = sort_k[1].get();
		r1.vals[0] = sort_v[0].get(), r1.vals[1] = sort_v[1].get();
		r1.hist = sort_h.get();
		radix_sort_pairs(r1, nh, (int)prim_bits, s);
		const int a1 = r1.result, b1 = a1 ^ 1;
		k_prp_keys<<<div_up(nh, 256), 256, 0, s>>>(nh, hits.get(), sort_v[a1].get(), prim_bits, sort_k[b1].get(),
			sort_v[b1].get());
		RadixSortBuffers rb;
		rb.keys[0] = sort_k[b1].get(), rb.keys[1] = sort_k[a1].get();
		rb.vals[0] = sort_v[b1].get(), rb.vals[1] = sort_v[a1].get();
		rb.hist = sort_h.get();
		const uint32_t slot_need = 32 - __builtin_clz(std::max(ns, 2u) - 1);
		radix_sort_pairs(rb, nh, (int)(slot_need + 2 + prim_bits), s);
		const uint32_t* keys = rb.keys[rb.result];
		const uint32_t* perm = rb.vals[rb.result];
		k_prp_gather<<<div_up(nh, 256), 256, 0, s>>>(nh, hits.get(), perm, hits_sorted.get(), hit_dst.get(), ns,
			slot_range.get());
		uint32_t* parent = comp_info.get();
		uint32_t* attached = parent + hit_cap;
		uint32_t* cnt = attached + hit_cap;
		uint32_t* cov = cnt + hit_cap;
		k_prp_uf_init<<<div_up(nh, 256), 256, 0, s>>>(nh, parent, attached, cnt, comp_key.get(), cov);
		// bucket table (2 entries per hit, generation-tagged; binary search beyond 2^24 hits)
		const uint64_t* table = nullptr;
		uint32_t		tcap = 1;
		uint64_t		tgen = 0;
		if (nh < (1u << 24))
		{
			while (tcap < 2u * nh)
				tcap <<= 1;
			if (tcap > bucket_cap || bucket_gen >= 255u)
			{
				bucket_cap = std::max(tcap, bucket_cap);
				buckets.reserve(bucket_cap);
				k_prp_fill_u64<<<div_up(bucket_cap, 256), 256, 0, s>>>(bucket_cap, buckets.get(), kHashEmpty);
				bucket_gen = 0;
			}
			tcap = bucket_cap;
			tgen = bucket_gen++;
			k_prp_bucket_insert<<<div_up(nh, 256), 256, 0, s>>>(nh, keys, buckets.get(), tcap, tgen);
			table = buckets.get();
		}
		HitArgs ha{ nh, V, prim_bits, hits_sorted.get(), keys, hit_dst.get(), slot_range.get(), table, tcap, tgen, parent,
			vv_off.get(), vv.get(), ve_off.get(), ve.get(), vf_off.get(), vf.get(), edges, faces, fae, eaf };
		k_prp_hit_union<<<div_up(nh, 256), 256, 0, s>>>(ha);
		k_prp_uf_compress<<<div_up(nh, 256), 256, 0, s>>>(nh, parent);
		// support hash: at most 6 entries per hit, load factor <= 1/2
		uint32_t hc = 1;
		while (hc < 12u * nh)
			hc <<= 1;
		if (hc > hash_cap || hash_gen >= 255u)
		{
			// new table or wrapped generation: clear to ~0 (generation 255, never used)
			hash_cap = std::max(hc, hash_cap);
			hash.reserve(hash_cap);
			k_prp_fill_u64<<<div_up(hash_cap, 256), 256, 0, s>>>(hash_cap, hash.get(), kHashEmpty);
			hash_gen = 0;
		}
		const uint64_t gen = hash_gen++;
		UT_REQUIRE(nh < (1u << 29), "prp: too many hits for the support keys");
		k_prp_fill_u32<<<1, 64, 0, s>>>(1, counters.get() + 1, 0u);
		UT_LAUNCH_CHECK();
		CompArgs ca{ nh, ns, V, batch0, hits_sorted.get(), parent, slots.get(), flag_off.get(), flags.get(), edges,
			faces, attached, cnt, comp_key.get(), cov, roots.get(), counters.get() + 1, hash.get(), hash_cap, gen,
			slot_best.get(), slot_root.get(), cpair_off.get(), cpair_rec.get() };
		k_prp_components<<<div_up(nh, 256), 256, 0, s>>>(ca, hits_sorted.get());
		k_prp_support<<<div_up(nh, 256), 256, 0, s>>>(ca);
		k_prp_coverage<<<1024, 256, 0, s>>>(ca);
		// key stage: slot_best[2j + 1] starts at the largest key
		k_prp_init_keys<<<div_up(ns, 256), 256, 0, s>>>(ns, slot_best.get());
		k_prp_select_key<<<div_up(nh, 256), 256, 0, s>>>(ca);
		k_prp_select_root<<<div_up(nh, 256), 256, 0, s>>>(ca);
		UT_LAUNCH_CHECK();
	}
	float* fn = normals.get();
	float* en = fn + 3 * (size_t)nf;
	float* vn = en + 3 * (size_t)ne;
	EvalArgs ea_{ ns, V, batch0, hits_sorted.get(), slot_range.get(), slot_root.get(),
		comp_info.get() + 3 * (size_t)hit_cap, slots.get(), cpair_off.get(), edges, faces, hop_b, va, ea, fa, mass,
		fn, en, vn, summary.get() };
	eval_part.reserve(sizeof(EvalPart) * (size_t)ns * kEvalSplit);
	EvalPart* ep = reinterpret_cast<EvalPart*>(eval_part.get());
	k_prp_eval<<<dim3(ns, kEvalSplit), 256, 0, s>>>(ea_, ep);
	k_prp_eval_final<<<div_up(ns, 64), 64, 0, s>>>(ea_, ep);
	UT_LAUNCH_CHECK();
	UT_HIP_CHECK(hipMemcpyAsync(out.data(), summary.get(), sizeof(PrpSummary) * ns, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	return nh;
}

uint32_t PrpResolver::resolve(const float* x, const ContourBuilder& ct, Lbvh& face_tree, Lbvh& edge_tree,
	std::vector<PrpChoice>& choices, PrpStats* stats)
{
	using clk = std::chrono::steady_clock;
	const auto		  t0 = clk::now();
	const hipStream_t s = ctx->stream;
	PrpStats		  st;
	ray_count[0] = ray_count[1] = ray_count[2] = 0;
	ee_prof = PrpKernelProfile{};
	prepare(x, ct, 0, 0);
	choices.assign(nc, PrpChoice{});
	const uint32_t active = std::min(nc, params.max_contours);
	st.contours = active;
	final_hits.reserve(std::max<size_t>(1024, final_cap));
	final_contour.reserve(std::max<size_t>(1024, final_cap));
	if (final_cap < 1024)
		final_cap = 1024;
	uint32_t n_final = 0;
	// hop and rest-geodesic distances per batch, searched only as far as the
	// k-rings reach (the thresholds of the distance kernel)
	if (active)
		dist_reserve(std::min(active, params.batch_contours));
	const auto t1 = clk::now();
	double	   ms_levels = 0.0, ms_final = 0.0;
	for (uint32_t c0 = 0; c0 < active; c0 += params.batch_contours)
	{
		const uint32_t n = std::min(params.batch_contours, active - c0);
		batch0 = c0, batch_n = n;
		const auto tl0 = clk::now();
		begin_batch(0);
		directions(c0, n);
		// distances out to each slot's first k-ring: hop threshold max(3, ceil(sqrt(boundary))),
		// the largest k that initial_k can choose.  A slot's diameter is known once its
		// search died out within the threshold (largest hop <= threshold); until then it is
		// above the threshold, and D stands for it as "infinite" in every comparison below.
		std::vector<uint32_t> kthr(2 * n), D(2 * n), req;
		std::vector<uint8_t>  dknown(2 * n);
		for (uint32_t sl = 0; sl < 2 * n; ++sl)
		{
			const uint32_t nb = h_b_off[2 * c0 + sl + 1] - h_b_off[2 * c0 + sl];
			kthr[sl] = std::max(3u, (uint32_t)std::ceil(std::sqrt((float)nb)));
			req.insert(req.end(), { sl, kthr[sl], 1u });
		}
		dist_pass(c0, req, false);
		dist_results(n);
		auto learn = [&](uint32_t sl)
		{
			dknown[sl] = h_dist_out[2 * sl] <= kthr[sl];
			D[sl] = dknown[sl] ? h_dist_out[2 * sl] : UINT32_MAX;
		};
		for (uint32_t sl = 0; sl < 2 * n; ++sl)
			learn(sl);
		// geodesic radii covering both roles' first rings
		req.clear();
		for (uint32_t sl = 0; sl < 2 * n; ++sl)
			req.insert(req.end(), { sl, 0u, 0u });
		dist_pass(c0, req, true);
		// adaptive k-ring driver (evaluate_current_runtime_directions_template :1058-1230)
		std::vector<uint32_t>				 ks(n), kd(n);
		std::vector<std::vector<ComboState>> R(n);
		std::vector<uint8_t>				 live(n, 1), grow(2 * n), extended(n);
		for (uint32_t lc = 0; lc < n; ++lc)
		{
			const uint32_t c = c0 + lc;
			ks[lc] = initial_k(h_b_off[2 * c + 1] - h_b_off[2 * c], D[2 * lc]);
			kd[lc] = initial_k(h_b_off[2 * c + 2] - h_b_off[2 * c + 1], D[2 * lc + 1]);
			R[lc].assign(h_dirs[c].size() / 3, ComboState{});
			choices[c].levels = 0;
		}
		for (uint32_t level = 0; level < 2 * 32 + 1; ++level)
		{
			std::vector<PrpSlot> sl;
			for (uint32_t lc = 0; lc < n; ++lc)
				if (live[lc])
					for (uint32_t ci = 0; ci < R[lc].size(); ++ci)
						if (!R[lc][ci].converged && !R[lc][ci].pruned)
						{
							const float* d = &h_dirs[c0 + lc][3 * ci];
							sl.push_back(PrpSlot{ lc, ci, ks[lc], kd[lc], { d[0], d[1], d[2] }, 0.0f, 0.0f, 0u });
						}
			if (sl.empty())
				break;
			std::vector<uint32_t> kr(2 * n);
			for (uint32_t lc = 0; lc < n; ++lc)
				kr[2 * lc] = ks[lc], kr[2 * lc + 1] = kd[lc];
			std::vector<PrpSummary> res;
			// evaluate in chunks that fit the slot key bits
			for (size_t b = 0; b < sl.size(); b += 4096)
			{
				std::vector<PrpSlot>	part(sl.begin() + b, sl.begin() + std::min(sl.size(), b + 4096));
				std::vector<PrpSummary> pr;
				st.hits += eval_slots(part, face_tree, edge_tree, kr, pr);
				res.insert(res.end(), pr.begin(), pr.end());
			}
			st.slots += (uint32_t)sl.size();
			++st.levels;
			for (size_t j = 0; j < sl.size(); ++j)
			{
				const uint32_t lc = sl[j].contour, ci = sl[j].combo;
				ComboState&	   r = R[lc][ci];
				r.present = true;
				r.s = res[j];
				r.ks = ks[lc], r.kd = kd[lc];
				const uint32_t D0 = D[2 * lc], D1 = D[2 * lc + 1];
				const bool	   full = ks[lc] >= D0 && kd[lc] >= D1;
				const bool	   src_ok = r.s.max_hop_src < ks[lc] || ks[lc] >= D0;
				const bool	   dst_ok = r.s.max_hop_dst < kd[lc] || kd[lc] >= D1;
				r.eligible = r.s.usable && src_ok && dst_ok;
				r.converged = r.eligible || full;
			}
			for (uint32_t lc = 0; lc < n; ++lc)
			{
				if (!live[lc])
					continue;
				choices[c0 + lc].levels = level + 1;
				auto& rr = R[lc];
				bool  has = false;
				float best = INFINITY;
				for (auto& r : rr)
					if (r.eligible && r.s.usable)
						has = true, best = std::min(best, response_objective(r.s));
				if (has)
					for (auto& r : rr)
						if (!r.converged && r.present && response_objective(r.s) >= best)
							r.pruned = true;
				bool		   pending = false, gs = false, gd = false;
				const uint32_t D0 = D[2 * lc], D1 = D[2 * lc + 1];
				for (auto& r : rr)
				{
					if (r.converged || r.pruned)
						continue;
					pending = true;
					gs = gs || (ks[lc] < D0 && (!r.s.usable || r.s.max_hop_src >= ks[lc]));
					gd = gd || (kd[lc] < D1 && (!r.s.usable || r.s.max_hop_dst >= kd[lc]));
				}
				if (!pending)
				{
					live[lc] = 0;
					continue;
				}
				grow[2 * lc] = gs, grow[2 * lc + 1] = gd;
			}
			// a growing ring whose diameter is not known yet: search out to the candidate
			// k first, so that grow_k sees the diameter whenever it is below the candidate
			req.clear();
			std::fill(extended.begin(), extended.end(), 0);
			for (uint32_t lc = 0; lc < n; ++lc)
				for (uint32_t r = 0; r < 2 && live[lc]; ++r)
				{
					const uint32_t sl = 2 * lc + r, k = r ? kd[lc] : ks[lc];
					if (grow[sl] && !dknown[sl] && 2 * k > kthr[sl])
					{
						kthr[sl] = 2 * k;
						req.insert(req.end(), { sl, kthr[sl], 0u });
						extended[lc] = 1;
					}
				}
			if (!req.empty())
			{
				dist_pass(c0, req, false);
				dist_results(n);
				for (size_t q = 0; q < req.size(); q += 3)
					learn(req[q]);
			}
			req.clear();
			for (uint32_t lc = 0; lc < n; ++lc)
			{
				if (!live[lc])
					continue;
				const uint32_t ns_ = grow[2 * lc] ? grow_k(ks[lc], D[2 * lc]) : ks[lc];
				const uint32_t nd_ = grow[2 * lc + 1] ? grow_k(kd[lc], D[2 * lc + 1]) : kd[lc];
				grow[2 * lc] = grow[2 * lc + 1] = 0;
				if (ns_ <= ks[lc] && nd_ <= kd[lc])
				{
					live[lc] = 0; // cannot grow further (the reference raises here)
					continue;
				}
				ks[lc] = ns_, kd[lc] = nd_;
				if (extended[lc])
					req.insert(req.end(), { 2 * lc, 0u, 0u, 2 * lc + 1, 0u, 0u });
			}
			// geodesic radii covering both roles' grown rings (a ring that grew inside an
			// unchanged hop threshold is covered by the radii set for that threshold)
			dist_pass(c0, req, true);
		}
		// selection (prp_select_response)
		std::vector<PrpSlot>  chosen;
		std::vector<uint32_t> kr(2 * n);
		for (uint32_t lc = 0; lc < n; ++lc)
		{
			int sel = -1;
			for (uint32_t ci = 0; ci < R[lc].size(); ++ci)
			{
				const ComboState& r = R[lc][ci];
				if (r.eligible && r.s.usable && (sel < 0 || better(r.s, ci, R[lc][sel].s, (uint32_t)sel)))
					sel = (int)ci;
			}
			PrpChoice& ch = choices[c0 + lc];
			kr[2 * lc] = ks[lc], kr[2 * lc + 1] = kd[lc];
			if (sel < 0)
				continue;
			const ComboState& r = R[lc][sel];
			const float*	  d = &h_dirs[c0 + lc][3 * sel];
			ch.combo = sel, ch.ks = r.ks, ch.kd = r.kd, ch.hit_count = r.s.hit_count, ch.covered = r.s.covered;
			ch.ef_count = r.s.ef_count, ch.depth = r.s.depth, ch.objective = r.s.objective, ch.area = r.s.area;
			ch.disp_cost = r.s.disp_cost;
			std::memcpy(ch.dir, d, 12);
			kr[2 * lc] = r.ks, kr[2 * lc + 1] = r.kd;
			chosen.push_back(PrpSlot{ lc, (uint32_t)sel, r.ks, r.kd, { d[0], d[1], d[2] }, 0.0f, 0.0f, 0u });
		}
		const auto tl1 = clk::now();
		ms_levels += std::chrono::duration<double, std::milli>(tl1 - tl0).count();
		if (chosen.empty())
			continue;
		// final pass: the kept hits of every chosen slot
		std::vector<PrpSummary> fin;
		const uint32_t			nh = eval_slots(chosen, face_tree, edge_tree, kr, fin);
		if (n_final + nh > final_cap)
		{
			// grow, keeping what is there
			const uint32_t		 ncap = std::max(2 * final_cap, n_final + nh + nh / 4);
			DevBuf<PrpHit>		 nfh;
			DevBuf<uint32_t>	 nfc;
			nfh.reserve(ncap), nfc.reserve(ncap);
			if (n_final)
			{
				UT_HIP_CHECK(hipMemcpyAsync(nfh.get(), final_hits.get(), sizeof(PrpHit) * n_final,
					hipMemcpyDeviceToDevice, s));
				UT_HIP_CHECK(hipMemcpyAsync(nfc.get(), final_contour.get(), 4 * (size_t)n_final,
					hipMemcpyDeviceToDevice, s));
			}
			UT_HIP_CHECK(hipStreamSynchronize(s));
			std::swap(final_hits.ptr, nfh.ptr), std::swap(final_hits.cap, nfh.cap);
			std::swap(final_contour.ptr, nfc.ptr), std::swap(final_contour.cap, nfc.cap);
			final_cap = ncap;
		}
		if (nh)
		{
			keep_pos.reserve(nh);
			scan_tmp.reserve(scan_tmp_size(nh));
			k_prp_keep_flags<<<div_up(nh, 256), 256, 0, s>>>(nh, hits_sorted.get(), slot_root.get(), keep_pos.get());
			exclusive_scan(keep_pos.get(), keep_pos.get(), nh, scan_tmp.get(), counters.get() + 2, s);
			k_prp_keep<<<div_up(nh, 256), 256, 0, s>>>(nh, hits_sorted.get(), slot_root.get(), slots.get(), c0,
				keep_pos.get(), n_final, final_hits.get(), final_contour.get(), final_cap);
			UT_LAUNCH_CHECK();
			uint32_t kept = 0;
			UT_HIP_CHECK(hipMemcpyAsync(&kept, counters.get() + 2, 4, hipMemcpyDeviceToHost, s));
			UT_HIP_CHECK(hipStreamSynchronize(s));
			n_final += kept;
		}
		ms_final += std::chrono::duration<double, std::milli>(clk::now() - tl1).count();
	}
	// contour suppression (:450-590) and responses
	const auto tf0 = clk::now();
	std::vector<uint8_t> sup(std::max(nc, 1u), 0);
	suppressed.reserve(sup.size());
	if (n_final && nc)
	{
		adjm.reserve((size_t)nc * nc);
		const uint32_t m = nc * nc;
		k_prp_fill_u8<<<div_up(m, 256), 256, 0, s>>>(m, adjm.get(), 0);
		k_prp_adjacent_contours<<<div_up(n_final, 256), 256, 0, s>>>(n_final, final_hits.get(), final_contour.get(),
			flag_off.get(), flags.get(), edges, faces, nc, adjm.get());
		UT_LAUNCH_CHECK();
		std::vector<uint8_t> A((size_t)nc * nc);
		UT_HIP_CHECK(hipMemcpyAsync(A.data(), adjm.get(), A.size(), hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		std::vector<uint32_t> order;
		for (uint32_t c = 0; c < active; ++c)
		{
			bool any = false;
			for (uint32_t o = 0; o < nc; ++o)
				any = any || (o != c && A[(size_t)c * nc + o]);
			if (any)
				order.push_back(c);
		}
		std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b)
			{
				if (choices[a].area != choices[b].area)
					return choices[a].area > choices[b].area;
				return a < b;
			});
		for (uint32_t o : order)
		{
			if (sup[o])
				continue;
			for (uint32_t a = 0; a < active; ++a)
				if (a != o && A[(size_t)o * nc + a])
					sup[a] = 1;
		}
		std::vector<float> cdir(3 * (size_t)nc, 0.0f), cC((size_t)nc, 0.0f);
		for (uint32_t c = 0; c < nc; ++c)
		{
			choices[c].suppressed = sup[c];
			std::memcpy(&cdir[3 * (size_t)c], choices[c].dir, 12);
			cC[c] = std::max(-choices[c].depth - 3e-3f, -params.response_depth);
		}
		upload(contour_c, cC, s);
		upload(contour_dir, cdir, s);
		upload(suppressed, sup, s);
		responses.reserve(n_final);
		response_meta.reserve(n_final);
		keep_pos.reserve(n_final);
		scan_tmp.reserve(scan_tmp_size(n_final));
		k_prp_response_flags<<<div_up(n_final, 256), 256, 0, s>>>(n_final, final_contour.get(), suppressed.get(),
			keep_pos.get());
		exclusive_scan(keep_pos.get(), keep_pos.get(), n_final, scan_tmp.get(), counters.get() + 3, s);
		k_prp_responses<<<div_up(n_final, 256), 256, 0, s>>>(n_final, final_hits.get(), final_contour.get(),
			contour_dir.get(), contour_c.get(), suppressed.get(), edges, faces, va, ea, fa, params.kappa,
			keep_pos.get(), responses.get(), response_meta.get(), n_final);
		UT_LAUNCH_CHECK();
	}
	else
		for (uint32_t c = 0; c < nc; ++c)
			choices[c].suppressed = 0;
	uint32_t nresp = 0;
	if (n_final)
	{
		UT_HIP_CHECK(hipMemcpyAsync(&nresp, counters.get() + 3, 4, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
	}
	last_final_hits = n_final;
	st.responses = nresp;
	for (int t = 0; t < 3; ++t)
		st.rays[t] = ray_count[t];
	st.ee = ee_prof;
	st.ms_prepare = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
	st.ms_levels = (float)ms_levels;
	st.ms_final = (float)(ms_final + std::chrono::duration<double, std::milli>(clk::now() - tf0).count());
	if (stats)
		*stats = st;
	return nresp;
}

} // namespace ut
