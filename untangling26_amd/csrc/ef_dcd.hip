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

#include <algorithm>

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

// A proximity pair whose weighted point difference vanishes (|d|^2 < 1e-10,
// the thickness term is 0 in the reference) turns into EF pairs with zero
// gradient and stiffness.  VF (v, f): one per edge around v, edge bary (1, 0)
// when v is the edge's first vertex else (0, 1), face bary |w1|, |w2|.  EE
// (e1, e2): for each face around e2 without a vertex of e1, (e1, face) with
// e1's bary and the face corners on e2 weighted by e1's bary; then for each
// face around e1 without a vertex of e2, (e2, face) with e2's bary and the
// corners on e1 weighted by e1's bary (the reference weights both with e1).
__global__ void k_ef_from_proximity(EfFromProximityArgs a)
{
	const uint32_t n = min(*a.rec_count, a.rec_cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const uint32_t* r = a.recs + 16 * (size_t)i;
		const uint32_t	r2 = r[2], r3 = r[3];
		const bool		is_vf = (r2 & 0x80000000u) != 0, is_ee = !is_vf && (r3 & 0x80000000u) != 0;
		if (!is_vf && !is_ee)
			continue;
		const uint32_t q[4] = { r[0] & 0x7FFFFFFFu, r[1] & 0x7FFFFFFFu, r2 & 0x7FFFFFFFu, r3 & 0x7FFFFFFFu };
		const float	   b0 = __uint_as_float(r[10]), b1 = __uint_as_float(r[11]);
		float		   w[4];
		if (is_vf)
			w[0] = 1.0f, w[1] = -b0, w[2] = -b1, w[3] = b0 + b1 - 1.0f;
		else
			w[0] = b0, w[1] = 1.0f - b0, w[2] = -b1, w[3] = b1 - 1.0f;
		float d[3];
		for (int c = 0; c < 3; ++c)
			d[c] = w[0] * a.x[3 * q[0] + c] + w[1] * a.x[3 * q[1] + c] + w[2] * a.x[3 * q[2] + c]
				+ w[3] * a.x[3 * q[3] + c];
		const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		if (d2 >= 1e-10f)
			continue;
		const uint2 id = reinterpret_cast<const uint2*>(a.rec_ids)[i];
		auto		emit = [&](uint32_t ev0, uint32_t ev1, uint32_t fid, float eb, float fb0, float fb1, uint32_t eid)
		{
			const uint32_t s = atomicAdd(a.count, 1u);
			if (s >= a.max_pairs)
				return;
			EfPairRec o;
			o.idx[0] = ev0, o.idx[1] = ev1;
			o.idx[2] = a.faces[3 * fid], o.idx[3] = a.faces[3 * fid + 1], o.idx[4] = a.faces[3 * fid + 2];
			o.bary[0] = eb, o.bary[1] = fb0, o.bary[2] = fb1;
			o.grad[0] = o.grad[1] = o.grad[2] = 0.0f;
			o.stiff = 0.0f;
			reinterpret_cast<EfPairRec*>(a.pairs)[s] = o;
			reinterpret_cast<uint2*>(a.pair_ids)[s] = make_uint2(eid, fid);
		};
		if (is_vf)
		{
			const uint32_t v = q[0], f = id.y;
			const float	   fb0 = fabsf(w[1]), fb1 = fabsf(w[2]);
			for (uint32_t k = a.vert_adj_off[v]; k < a.vert_adj_off[v + 1]; ++k)
			{
				const uint32_t ae = a.vert_adj_edges[k];
				const uint32_t ev0 = a.edges[2 * ae], ev1 = a.edges[2 * ae + 1];
				emit(ev0, ev1, f, ev0 == v ? 1.0f : 0.0f, fb0, fb1, ae);
			}
			continue;
		}
		const uint32_t e1 = id.x, e2 = id.y;
		const uint32_t a0 = a.edges[2 * e1], a1 = a.edges[2 * e1 + 1], c0 = a.edges[2 * e2], c1 = a.edges[2 * e2 + 1];
		const float	   eb1[2] = { fabsf(w[0]), fabsf(w[1]) }, eb2x = fabsf(w[2]);
		for (int side = 0; side < 2; ++side)
		{
			const uint32_t fe = side ? e1 : e2, ee = side ? e2 : e1;
			const uint32_t s0 = side ? a0 : c0, s1 = side ? a1 : c1; // corners the face shares
			const uint32_t x0 = side ? c0 : a0, x1 = side ? c1 : a1; // must not be in the face
			for (int k = 0; k < 2; ++k)
			{
				const uint32_t af = a.edge_adj_faces[2 * fe + k];
				if (af == kNone)
					continue;
				const uint32_t F[3] = { a.faces[3 * af], a.faces[3 * af + 1], a.faces[3 * af + 2] };
				if (F[0] == x0 || F[1] == x0 || F[2] == x0 || F[0] == x1 || F[1] == x1 || F[2] == x1)
					continue;
				float fb[3] = { 0.0f, 0.0f, 0.0f };
				int	  matched = 0;
				for (int j = 0; j < 3; ++j)
				{
					if (F[j] == s0)
						fb[j] = eb1[0], ++matched;
					else if (F[j] == s1)
						fb[j] = eb1[1], ++matched;
				}
				if (matched == 2)
					emit(a.edges[2 * ee], a.edges[2 * ee + 1], af, side ? eb2x : eb1[0], fb[0], fb[1], ee);
			}
		}
	}
}

void ef_from_proximity(Context* ctx, const EfFromProximityArgs& a)
{
	UT_REQUIRE(a.recs && a.rec_ids && a.rec_count && a.x && a.edges && a.faces && a.vert_adj_off && a.vert_adj_edges
			&& a.edge_adj_faces && a.count,
		"ef_from_proximity: null input");
	UT_REQUIRE(a.max_pairs == 0 || (a.pairs && a.pair_ids), "ef_from_proximity: null output");
	if (a.rec_cap == 0)
		return;
	const uint32_t blocks = std::min<uint32_t>(div_up(a.rec_cap, 256), 2048);
	k_ef_from_proximity<<<blocks, 256, 0, ctx->stream>>>(a);
	UT_LAUNCH_CHECK();
}

} // namespace ut
