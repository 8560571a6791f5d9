// Edge-face intersection detection (DCD) for gfx950.
//
// fn_narrow_phase_ef_dcd_query, intersection_resolver.cpp:86-197, with the
// segment/triangle test LineIntersection (intersection_resolver_helper.cpp:
// 333-481).  All arithmetic is FP32 with no contraction so the result set is
// reproducible bit for bit on any IEEE host.
#include "ut_common.hpp"
#include "ut_runtime.hpp"
#include "ef_dcd.hpp"
#include "lbvh_traverse.hpp"
#include "seg_tri.hpp"

namespace ut
{

// Returns true and fills the pair when edge `eid` properly cuts face `fid`.
__device__ __forceinline__ bool ef_test(uint32_t eid,
	uint32_t						   fid,
	const float* __restrict__ x,
	const uint32_t* __restrict__ edges,
	const uint32_t* __restrict__ faces,
	const uint32_t* __restrict__ vprop,
	const float* __restrict__ d_hat,
	EfPairRec* out)
{
	const uint32_t e0 = edges[2 * eid], e1 = edges[2 * eid + 1];
	const uint32_t f0 = faces[3 * fid], f1 = faces[3 * fid + 1], f2 = faces[3 * fid + 2];
	if (e0 == f0 || e0 == f1 || e0 == f2 || e1 == f0 || e1 == f1 || e1 == f2)
		return false;
	if (vprop)
	{
		const uint32_t pl = vprop[e0], pr = vprop[f0];
		if ((pl & kPropFixed) && (pr & kPropFixed))
			return false;
		if ((pl & kPropRigid) && (pr & kPropRigid) && (pl >> kPropObjectShift) == (pr >> kPropObjectShift))
			return false;
	}
	if (d_hat[e0] + d_hat[f0] == 0.0f)
		return false;
	const f3 P0 = ld3(x, e0), P1 = ld3(x, e1);
	const f3 A = ld3(x, f0), B = ld3(x, f1), C = ld3(x, f2);
	f3		 bary;
	float	 t;
	if (!line_intersection(P0, P1, A, B, C, &bary, &t))
		return false;
	// Grazing filters (intersection_resolver.cpp:153-179).
	const f3	nun = cross3(B - A, C - A);
	const float nlen = sqrtf(dot3(nun, nun));
	const float face_char = sqrtf(fmaxf(nlen, 1e-30f));
	const float nl = fmaxf(nlen, 1e-30f);
	const f3	n = mk3(nun.x / nl, nun.y / nl, nun.z / nl);
	const float d0 = dot3(P0 - A, n);
	const float d1 = dot3(P1 - A, n);
	const float pen = fminf(fabsf(d0), fabsf(d1));
	if (pen < 1e-4f * face_char)
		return false;
	if (fminf(fminf(bary.x, bary.y), bary.z) < 1e-4f)
		return false;
	if (fminf(t, 1.0f - t) < 1e-4f)
		return false;
	out->idx[0] = e0;
	out->idx[1] = e1;
	out->idx[2] = f0;
	out->idx[3] = f1;
	out->idx[4] = f2;
	out->bary[0] = 1.0f - t;
	out->bary[1] = bary.x;
	out->bary[2] = bary.y;
	out->grad[0] = out->grad[1] = out->grad[2] = 0.0f;
	out->stiff = 0.0f;
	return true;
}

__device__ __forceinline__ void ef_emit(bool hit, const EfPairRec& rec, uint32_t eid, uint32_t fid, uint32_t* count,
	EfPairRec* pairs, uint2* pair_ids, uint32_t max_pairs)
{
	const uint32_t slot = wave_append(count, hit);
	if (hit && slot < max_pairs)
	{
		pairs[slot] = rec;
		pair_ids[slot] = make_uint2(eid, fid);
	}
}

__global__ __launch_bounds__(256) void k_ef_dcd(const uint2* __restrict__ broad,
	const uint32_t* __restrict__ broad_count,
	uint32_t broad_cap,
	const float* __restrict__ x,
	const uint32_t* __restrict__ edges,
	const uint32_t* __restrict__ faces,
	const uint32_t* __restrict__ vprop,
	const float* __restrict__ d_hat,
	uint32_t* __restrict__ count,
	EfPairRec* __restrict__ pairs,
	uint2* __restrict__ pair_ids,
	uint32_t max_pairs)
{
	const uint32_t nb = min(*broad_count, broad_cap);
	// grid-stride so the launch size does not depend on the device-side count
	for (uint32_t base = blockIdx.x * blockDim.x; base < nb; base += gridDim.x * blockDim.x)
	{
		const uint32_t i = base + threadIdx.x;
		EfPairRec	   rec;
		bool		   hit = false;
		uint2		   p = make_uint2(0, 0);
		if (i < nb)
		{
			p = broad[i];
			hit = ef_test(p.x, p.y, x, edges, faces, vprop, d_hat, &rec);
		}
		ef_emit(hit, rec, p.x, p.y, count, pairs, pair_ids, max_pairs);
	}
}

// Fused broad + narrow phase: unculled edge-vs-face traversal whose staged
// candidates go straight into the EF test, so the broad list is never written.
struct EfNarrow
{
	const float*	x;
	const uint32_t* edges;
	const uint32_t* faces;
	const uint32_t* vprop;
	const float*	d_hat;
	uint32_t*		count;
	EfPairRec*		pairs;
	uint2*			pair_ids;
	uint32_t		max_pairs;
	__device__ __forceinline__ void operator()(bool valid, uint32_t eid, uint32_t fid)
	{
		EfPairRec rec;
		const bool hit = valid && ef_test(eid, fid, x, edges, faces, vprop, d_hat, &rec);
		ef_emit(hit, rec, eid, fid, count, pairs, pair_ids, max_pairs);
	}
};

__global__ __launch_bounds__(kTravThreads) void k_ef_detect(uint32_t ne,
	const float* __restrict__ x,
	const uint32_t* __restrict__ edges,
	const uint32_t* __restrict__ faces,
	const uint32_t* __restrict__ vprop,
	const float* __restrict__ d_hat,
	const float* __restrict__ thickness,
	const BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ status,
	uint32_t* __restrict__ spill,
	uint32_t max_visits,
	uint32_t* __restrict__ count,
	EfPairRec* __restrict__ pairs,
	uint2* __restrict__ pair_ids,
	uint32_t max_pairs)
{
	__shared__ TravShared sh;
	const uint32_t		  q = blockIdx.x * kTravThreads + threadIdx.x;
	const bool			  active = q < ne;
	const box3 qb = active ? query_box(kQueryEdgesUnculled, q, x, x, edges, d_hat, thickness) : box_empty();
	EfNarrow   nar{ x, edges, faces, vprop, d_hat, count, pairs, pair_ids, max_pairs };
	traverse_staged(sh, active, q, qb, false, nodes, status, spill, max_visits, nar);
}

void ef_detect(Context* ctx, Lbvh& face_tree, const EfDcdArgs& a, uint32_t num_edges, const float* thickness)
{
	UT_REQUIRE(face_tree.type == kTreeFace, "ef_detect needs the face tree");
	UT_REQUIRE(a.x && a.edges && a.faces && a.d_hat && a.count && thickness, "ef_detect: null input");
	UT_REQUIRE(a.max_pairs == 0 || (a.pairs && a.pair_ids), "ef_detect: null output");
	if (num_edges == 0)
		return;
	face_tree.spill.reserve((size_t)num_edges * (kStackTotal - kStackLds));
	k_ef_detect<<<div_up(num_edges, kTravThreads), kTravThreads, 0, ctx->stream>>>(num_edges,
		a.x,
		a.edges,
		a.faces,
		a.vprop,
		a.d_hat,
		thickness,
		face_tree.nodes.get(),
		face_tree.status.get(),
		face_tree.spill.get(),
		face_tree.num_leaves + 1,
		a.count,
		reinterpret_cast<EfPairRec*>(a.pairs),
		reinterpret_cast<uint2*>(a.pair_ids),
		a.max_pairs);
	UT_LAUNCH_CHECK();
}

void ef_dcd(Context* ctx, const EfDcdArgs& a)
{
	UT_REQUIRE(a.broad && a.broad_count && a.x && a.edges && a.faces && a.d_hat && a.count, "ef_dcd: null input");
	UT_REQUIRE(a.max_pairs == 0 || (a.pairs && a.pair_ids), "ef_dcd: null output");
	int dev_cu = 256;
	const uint32_t blocks = (uint32_t)dev_cu * 8;
	k_ef_dcd<<<blocks, 256, 0, ctx->stream>>>(reinterpret_cast<const uint2*>(a.broad),
		a.broad_count,
		a.broad_cap,
		a.x,
		a.edges,
		a.faces,
		a.vprop,
		a.d_hat,
		a.count,
		reinterpret_cast<EfPairRec*>(a.pairs),
		reinterpret_cast<uint2*>(a.pair_ids),
		a.max_pairs);
	UT_LAUNCH_CHECK();
}

} // namespace ut
