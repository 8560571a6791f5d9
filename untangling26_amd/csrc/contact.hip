// Proximity DCD (VF / EE with the IPC barrier) and additive CCD on gfx950.
//
// fn_narrow_phase_{vf,ee}_dcd_query (narrow_phase.cpp:980-1346) and
// fn_narrow_phase_{vf,ee}_ccd_query (narrow_phase.cpp:412-800).  Distances,
// barrier derivatives and the time of impact are evaluated in FP64 (the
// reference runs them in FP32 fast-math); the emitted CollisionPairTemplate
// keeps the reference's FP32 fields.
//
// Each query comes in two forms: over a broad-phase pair list (the
// reference's two steps) and fused with the LBVH traversal, where the staged
// candidate batches of lbvh_traverse.hpp feed the narrow test directly.
#include "ut_common.hpp"
#include "contact.hpp"
#include "contact_math.hpp"
#include "friction_math.hpp"
#include "lbvh_traverse.hpp"
#include "radix_sort.hpp"

#include <algorithm>

namespace ut
{

constexpr double kMinDist = 5e-6;						 // min_distance_threshold, narrow_phase.cpp:959
constexpr double kMinDistSqr = kMinDist * kMinDist;		 // narrow_phase.cpp:960
constexpr uint32_t kTypeMask = 0x80000000u;			 // mask_collision_type, collision_data.h:36

__device__ __forceinline__ bool invalid_combo(const uint32_t* vprop, uint32_t a, uint32_t b)
{
	if (!vprop)
		return false;
	const uint32_t pa = vprop[a], pb = vprop[b];
	return ((pa & kPropFixed) && (pb & kPropFixed))
		|| ((pa & kPropRigid) && (pb & kPropRigid) && (pa >> kPropObjectShift) == (pb >> kPropObjectShift));
}

// stiffness pair (k1, k2) of narrow_phase.cpp:1091-1112
__device__ __forceinline__ void contact_stiffness(int energy, double d, double thickness, double d_hat, double area,
	double kappa, double* k1, double* k2)
{
	if (energy == kEnergyQuadratic)
	{
		const double C = (d - thickness) - d_hat;
		const double stiff = area * kappa;
		*k1 = stiff * C;
		*k2 = stiff;
	}
	else
	{
		*k1 = area * kappa * barrier_d1(d - thickness, d_hat);
		*k2 = area * kappa * barrier_d2(d - thickness, d_hat);
	}
}

__device__ __forceinline__ bool vf_dcd_test(uint32_t vid, uint32_t fid, const ContactArgs& a, ContactRec* out)
{
	const uint32_t f0 = a.faces[3 * fid], f1 = a.faces[3 * fid + 1], f2 = a.faces[3 * fid + 2];
	if (vid == f0 || vid == f1 || vid == f2 || invalid_combo(a.vprop, vid, f0))
		return false;
	const v3<double> p = load_v3<double>(a.x, vid);
	const v3<double> t0 = load_v3<double>(a.x, f0), t1 = load_v3<double>(a.x, f1), t2 = load_v3<double>(a.x, f2);
	const v3<double> b = pt_coeff(p, t0, t1, t2);
	const v3<double> diff = b.x * (p - t0) + b.y * (p - t1) + b.z * (p - t2);
	const double	 d2 = dot(diff, diff);
	const double	 d_hat = ((double)a.d_hat[vid] + (double)a.d_hat[f0]) * 0.5;
	const double	 th = (double)a.offset[vid] + (double)a.offset[f0];
	if (!(d2 < (th + d_hat) * (th + d_hat) && (d2 > kMinDistSqr || a.energy == kEnergyQuadratic)))
		return false;
	const double d = sqrt(d2);
	v3<double>	 n = V3m(diff.x / d, diff.y / d, diff.z / d);
	if (d < kMinDistSqr) // the reference compares d with the squared threshold
		n = V3m(0.0, 1.0, 0.0);
	const double area = 0.5 * ((double)a.rest_vert_area[vid] + (double)a.rest_face_area[fid]);
	double		 k1, k2;
	contact_stiffness(a.energy, d, th, d_hat, area, a.kappa, &k1, &k2);
	if (!isfinite(k1) || !isfinite(k2))
		return false;
	out->idx[0] = vid;
	out->idx[1] = f0;
	out->idx[2] = f1 | kTypeMask;
	out->idx[3] = f2;
	out->normal[0] = (float)n.x;
	out->normal[1] = (float)n.y;
	out->normal[2] = (float)n.z;
	out->area = (float)area;
	out->k1 = (float)k1;
	out->k2 = (float)k2;
	out->bary[0] = (float)b.x;
	out->bary[1] = (float)b.y;
	out->dv[0] = out->dv[1] = out->dv[2] = 0.0f;
	out->lambda = 0.0f;
	return true;
}

__device__ __forceinline__ bool ee_dcd_test(uint32_t left, uint32_t right, const ContactArgs& a, ContactRec* out)
{
	const uint32_t l0 = a.edges[2 * left], l1 = a.edges[2 * left + 1];
	const uint32_t r0 = a.edges[2 * right], r1 = a.edges[2 * right + 1];
	if (l0 == r0 || l0 == r1 || l1 == r0 || l1 == r1 || invalid_combo(a.vprop, l0, r0))
		return false;
	const v3<double> a0 = load_v3<double>(a.x, l0), a1 = load_v3<double>(a.x, l1);
	const v3<double> b0 = load_v3<double>(a.x, r0), b1 = load_v3<double>(a.x, r1);
	const c4<double> c = ee_coeff(a0, a1, b0, b1);
	const v3<double> x0 = c.a * a0 + c.b * a1;
	const v3<double> x1 = c.c * b0 + c.d * b1;
	const v3<double> diff = x0 - x1;
	const double	 d2 = dot(diff, diff);
	const double	 d_hat = ((double)a.d_hat[l0] + (double)a.d_hat[r0]) * 0.5;
	const double	 th = (double)a.offset[l0] + (double)a.offset[r0];
	if (!(d2 < (th + d_hat) * (th + d_hat) && (d2 > kMinDistSqr || a.energy == kEnergyQuadratic)))
		return false;
	const double d = sqrt(d2);
	v3<double>	 n = V3m(diff.x / d, diff.y / d, diff.z / d);
	if (d < kMinDistSqr)
		n = V3m(0.0, 1.0, 0.0);
	const double area = 0.5 * ((double)a.rest_edge_area[left] + (double)a.rest_edge_area[right]);
	double		 k1, k2;
	contact_stiffness(a.energy, d, th, d_hat, area, a.kappa, &k1, &k2);
	if (!isfinite(k1) || !isfinite(k2))
		return false;
	out->idx[0] = l0;
	out->idx[1] = l1;
	out->idx[2] = r0;
	out->idx[3] = r1 | kTypeMask;
	out->normal[0] = (float)n.x;
	out->normal[1] = (float)n.y;
	out->normal[2] = (float)n.z;
	out->area = (float)area;
	out->k1 = (float)k1;
	out->k2 = (float)k2;
	out->bary[0] = (float)c.a;
	out->bary[1] = (float)c.c;
	out->dv[0] = out->dv[1] = out->dv[2] = 0.0f;
	out->lambda = 0.0f;
	return true;
}

__device__ __forceinline__ double vf_ccd(uint32_t vid, uint32_t fid, const ContactArgs& a)
{
	const uint32_t f0 = a.faces[3 * fid], f1 = a.faces[3 * fid + 1], f2 = a.faces[3 * fid + 2];
	if (vid == f0 || vid == f1 || vid == f2 || invalid_combo(a.vprop, vid, f0))
		return kLineSearchMaxT;
	v3<double> x0[4] = { load_v3<double>(a.x, vid), load_v3<double>(a.x, f0), load_v3<double>(a.x, f1),
		load_v3<double>(a.x, f2) };
	v3<double> dx[4] = { load_v3<double>(a.x_end, vid) - x0[0], load_v3<double>(a.x_end, f0) - x0[1],
		load_v3<double>(a.x_end, f1) - x0[2], load_v3<double>(a.x_end, f2) - x0[3] };
	const double th = (double)a.offset[vid] + (double)a.offset[f0];
	return toi_postprocess(accd_toi(false, x0, dx, th, a.toi));
}

__device__ __forceinline__ double ee_ccd(uint32_t left, uint32_t right, const ContactArgs& a)
{
	const uint32_t l0 = a.edges[2 * left], l1 = a.edges[2 * left + 1];
	const uint32_t r0 = a.edges[2 * right], r1 = a.edges[2 * right + 1];
	if (l0 == r0 || l0 == r1 || l1 == r0 || l1 == r1 || invalid_combo(a.vprop, l0, r0))
		return kLineSearchMaxT;
	v3<double> x0[4] = { load_v3<double>(a.x, l0), load_v3<double>(a.x, l1), load_v3<double>(a.x, r0),
		load_v3<double>(a.x, r1) };
	v3<double> dx[4] = { load_v3<double>(a.x_end, l0) - x0[0], load_v3<double>(a.x_end, l1) - x0[1],
		load_v3<double>(a.x_end, r0) - x0[2], load_v3<double>(a.x_end, r1) - x0[3] };
	const double th = (double)a.offset[l0] + (double)a.offset[r1]; // narrow_phase.cpp:656 uses right_edge[1]
	return toi_postprocess(accd_toi(true, x0, dx, th, a.toi));
}

// Non-negative floats order like their bit patterns.  The minimum only
// decreases, so a relaxed read that is not above t (stale or not) proves the
// atomic would change nothing: atomics on one word serialise, and most pairs
// with a time of impact do not lower the running minimum.
__device__ __forceinline__ void toi_min(float* toi, double t)
{
	uint32_t* const w = reinterpret_cast<uint32_t*>(toi);
	const uint32_t	b = __float_as_uint((float)t);
	if (b < __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
		atomicMin(w, b);
}

struct NarrowOp
{
	const ContactArgs* a;
	int				   kind; // 0 VF DCD, 1 EE DCD, 2 VF CCD, 3 EE CCD
	// fn_filter_broadphase_{vf,ee}_pairs (intersection_resolver.cpp:2427-2535)
	__device__ __forceinline__ bool ccd_filtered(uint32_t q, uint32_t leaf) const
	{
		if (!a->filter.off && !a->filter.invalid)
			return false;
		uint32_t s[3], t[3];
		int		 ns, nt;
		ccd_pair_sides(kind, q, leaf, a->edges, a->faces, s, &ns, t, &nt);
		return pair_filtered(a->filter, s, ns, t, nt);
	}
	__device__ __forceinline__ void operator()(bool valid, uint32_t q, uint32_t leaf)
	{
		if (kind <= 1)
		{
			ContactRec rec;
			const bool hit = valid && (kind == 0 ? vf_dcd_test(q, leaf, *a, &rec) : ee_dcd_test(q, leaf, *a, &rec));
			const uint32_t s = wave_append(a->count, hit);
			if (hit && s < a->max_pairs)
			{
				a->pairs[s] = rec;
				a->pair_ids[s] = make_uint2(q, leaf);
			}
		}
		else if (valid && !ccd_filtered(q, leaf))
		{
			const double t = kind == 2 ? vf_ccd(q, leaf, *a) : ee_ccd(q, leaf, *a);
			if (t != kLineSearchMaxT)
				toi_min(a->toi, t);
		}
	}
};

__global__ __launch_bounds__(kTravThreads) void k_contact_fused(int kind, ContactArgs a, const BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ status, uint32_t* __restrict__ spill, uint32_t max_visits)
{
	__shared__ TravShared sh;
	const bool			  vf = (kind == 0 || kind == 2);
	const uint32_t		  nq = vf ? a.num_verts : a.num_edges;
	const uint32_t		  q = blockIdx.x * kTravThreads + threadIdx.x;
	const bool			  active = q < nq;
	const float*		  xe = kind >= 2 ? a.x_end : a.x;
	const box3			  qb = active ? query_box(vf ? kQueryVerts : kQueryEdges, q, a.x, xe, a.edges, a.d_hat, a.offset)
										: box_empty();
	NarrowOp			  op{ &a, kind };
	traverse_staged(sh, active, q, qb, !vf, nodes, status, spill, max_visits, op);
}

// CCD broad phase: the swept-box traversal lists the (query, leaf) pairs that
// survive the region filter; the ACCD runs in its own kernel over that list, so
// the traversal is not held to the FP64 ACCD's register budget (occupancy).
// The list is kCcdLists sub-lists of `cap` pairs, one counter each (64 B
// apart): a single shared counter would serialise the waves' appends.
constexpr uint32_t kCcdLists = 64;
constexpr uint32_t kCcdListStride = 16;
struct CcdEmit
{
	const NarrowOp* op;
	uint32_t*		count; // [kCcdLists * kCcdListStride]
	uint2*			list;  // [kCcdLists * cap]
	uint32_t		cap;
	__device__ __forceinline__ void operator()(bool valid, uint32_t q, uint32_t leaf)
	{
		const uint32_t k = (blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) % kCcdLists;
		const bool	   keep = valid && !op->ccd_filtered(q, leaf);
		const uint32_t s = wave_append(&count[kCcdListStride * k], keep);
		if (keep && s < cap)
			list[(size_t)k * cap + s] = make_uint2(q, leaf);
	}
};

constexpr int kCcdStackLds = 16;

__global__ __launch_bounds__(kTravThreads) void k_ccd_broad(int kind, ContactArgs a, const BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ status, uint32_t* __restrict__ spill, uint32_t max_visits, uint32_t* __restrict__ count,
	uint2* __restrict__ list, uint32_t cap)
{
	__shared__ TravSharedT<kCcdStackLds> sh;
	const bool		vf = kind == 2;
	const uint32_t	nq = vf ? a.num_verts : a.num_edges;
	const uint32_t	q = blockIdx.x * kTravThreads + threadIdx.x;
	const bool		active = q < nq;
	const box3		qb = active ? query_box(vf ? kQueryVerts : kQueryEdges, q, a.x, a.x_end, a.edges, a.d_hat, a.offset)
									: box_empty();
	const NarrowOp	op{ &a, kind };
	CcdEmit			emit{ &op, count, list, cap };
	auto			test = [&](const box3& b) { return box_overlap(b, qb); };
	traverse_staged_test(sh, active, q, q, test, !vf, nodes, status, spill, max_visits, emit);
}

// ACCD over the listed pairs (blockIdx.y = sub-list); the running minimum TOI
// lets pairs stop early.
__global__ __launch_bounds__(256) void k_ccd_pairs(int kind, ContactArgs a, const uint2* __restrict__ list,
	const uint32_t* __restrict__ count, uint32_t cap)
{
	const uint32_t k = blockIdx.y;
	const uint32_t n = min(count[kCcdListStride * k], cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const uint2	 p = list[(size_t)k * cap + i];
		const double t = kind == 2 ? vf_ccd(p.x, p.y, a) : ee_ccd(p.x, p.y, a);
		if (t != kLineSearchMaxT)
			toi_min(a.toi, t);
	}
}

__global__ __launch_bounds__(256) void k_contact_list(int kind, ContactArgs a, const uint2* __restrict__ broad,
	const uint32_t* __restrict__ broad_count, uint32_t broad_cap)
{
	const uint32_t nb = min(*broad_count, broad_cap);
	NarrowOp	   op{ &a, kind };
	for (uint32_t base = blockIdx.x * blockDim.x; base < nb; base += gridDim.x * blockDim.x)
	{
		const uint32_t i = base + threadIdx.x;
		const bool	   valid = i < nb;
		const uint2	   p = valid ? broad[i] : make_uint2(0u, 0u);
		op(valid, p.x, p.y);
	}
}

__global__ void k_toi_init(float* toi) { *toi = (float)kLineSearchMaxT; }

void contact_query(Context* ctx, int kind, const ContactArgs& a, Lbvh* tree, const uint32_t* broad,
	const uint32_t* broad_count, uint32_t broad_cap)
{
	UT_REQUIRE(kind >= 0 && kind <= 3, "contact query: unknown kind");
	UT_REQUIRE(a.x && a.d_hat && a.offset, "contact query: null positions or per-vertex arrays");
	if (kind == 0 || kind == 2)
		UT_REQUIRE(a.faces, "VF query needs faces");
	else
		UT_REQUIRE(a.edges, "EE query needs edges");
	if (kind <= 1)
		UT_REQUIRE(a.count && (a.max_pairs == 0 || (a.pairs && a.pair_ids)) && a.rest_vert_area && a.rest_face_area
				&& a.rest_edge_area,
			"DCD query: null output or rest areas");
	else
		UT_REQUIRE(a.x_end && a.toi, "CCD query needs x_end and toi");
	if (tree)
	{
		const bool vf = (kind == 0 || kind == 2);
		UT_REQUIRE(tree->type == (vf ? kTreeFace : kTreeEdge), "fused query: wrong tree type");
		const uint32_t nq = vf ? a.num_verts : a.num_edges;
		if (nq == 0)
			return;
		// per-query spill region: the CCD listing keeps fewer stack entries in LDS
		tree->spill.reserve((size_t)nq * (kStackTotal - std::min(kStackLds, kCcdStackLds)));
		if (kind <= 1)
			k_contact_fused<<<div_up(nq, kTravThreads), kTravThreads, 0, ctx->stream>>>(
				kind, a, tree->nodes.get(), tree->status.get(), tree->spill.get(), tree->num_leaves + 1);
		else
		{
			// list the swept pairs (grow the lists and list again on overflow), then the ACCD
			tree->pair_count.reserve(kCcdLists * kCcdListStride);
			if (tree->pair_buf.cap < (1u << 21))
				tree->pair_buf.reserve(1u << 21);
			std::vector<uint32_t> cnt(kCcdLists * kCcdListStride);
			uint32_t			  cap = 0, mx = 0;
			for (int attempt = 0; attempt < 8; ++attempt)
			{
				cap = (uint32_t)(tree->pair_buf.cap / kCcdLists);
				UT_HIP_CHECK(hipMemsetAsync(tree->pair_count.get(), 0, 4 * cnt.size(), ctx->stream));
				k_ccd_broad<<<div_up(nq, kTravThreads), kTravThreads, 0, ctx->stream>>>(kind, a, tree->nodes.get(),
					tree->status.get(), tree->spill.get(), tree->num_leaves + 1, tree->pair_count.get(),
					tree->pair_buf.get(), cap);
				UT_LAUNCH_CHECK();
				UT_HIP_CHECK(
					hipMemcpyAsync(cnt.data(), tree->pair_count.get(), 4 * cnt.size(), hipMemcpyDeviceToHost, ctx->stream));
				UT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
				mx = 0;
				for (uint32_t k = 0; k < kCcdLists; ++k)
					mx = std::max(mx, cnt[kCcdListStride * k]);
				if (mx <= cap)
					break;
				tree->pair_buf.reserve((size_t)(mx + mx / 4) * kCcdLists);
			}
			UT_REQUIRE(mx <= cap, "CCD: swept pair lists could not be grown");
			if (mx)
				k_ccd_pairs<<<dim3(std::min<uint32_t>(div_up(mx, 256), 64), kCcdLists), 256, 0, ctx->stream>>>(
					kind, a, tree->pair_buf.get(), tree->pair_count.get(), cap);
		}
	}
	else
	{
		UT_REQUIRE(broad && broad_count, "list query needs the broad-phase list");
		k_contact_list<<<256 * 8, 256, 0, ctx->stream>>>(kind, a, reinterpret_cast<const uint2*>(broad), broad_count,
			broad_cap);
	}
	UT_LAUNCH_CHECK();
}

void toi_reset(Context* ctx, float* toi)
{
	k_toi_init<<<1, 1, 0, ctx->stream>>>(toi);
	UT_LAUNCH_CHECK();
}

namespace
{
constexpr int kEnergyThreads = 256;

// Contact part of the incremental potential at x (FP32 or FP64 positions P).
//
// Barrier term: for every proximity record the closest-point distance is
// recomputed in FP64 from x (pt / ee coefficients, as the DCD that produced the
// record), then kappa * area * b(d - th, d_hat) for the barrier or
// kappa * area * C^2 / 2, C = d - th - d_hat, for the quadratic energy
// (narrow_phase.cpp:1091-1112 integrates to these); its gradient
// kappa * area * b'(d - th) w_k n with the re-solved weights w and normal n
// (the closest point is optimal, so d's derivative is w_k n).  Records with
// k2 = 0 (disabled, or lagged friction pairs) carry no barrier energy.
//
// Friction term: for every record with a friction multiplier, the lagged IPC
// friction of the tangential displacement sum_k w_k (x_k - x_start_k) with the
// record's weights and normal (fn_compute_repulsion_energy's friction part,
// narrow_phase.cpp:3485-3512, and its gradient, friction_math.hpp).
template <typename P>
__global__ __launch_bounds__(kEnergyThreads) void k_contact_potential(const ContactRec* __restrict__ pairs,
	const uint32_t* __restrict__ count, uint32_t cap, const P* __restrict__ x, const P* __restrict__ x_start,
	const float* __restrict__ d_hat, const float* __restrict__ offset, double kappa, int energy, int terms,
	double* __restrict__ grad, double* __restrict__ part)
{
	const uint32_t n = min(*count, cap);
	double		   acc_b = 0.0, acc_f = 0.0;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const ContactRec r = pairs[i];
		uint32_t		 v[4];
		v3<double>		 p[4];
		for (int k = 0; k < 4; ++k)
		{
			v[k] = r.idx[k] & kIndexMask;
			p[k] = load_v3<double>(x, v[k]);
		}
		const bool vf = (r.idx[2] & kTypeBit) != 0u;
		if ((terms & kTermBarrier) && r.k2 != 0.0f)
		{
			double	 w[4];
			uint32_t ka, kb; // the two vertices whose d_hat / offset the DCD used
			if (vf)
			{
				const v3<double> b = pt_coeff(p[0], p[1], p[2], p[3]);
				w[0] = b.x + b.y + b.z, w[1] = -b.x, w[2] = -b.y, w[3] = -b.z;
				ka = v[0], kb = v[1];
			}
			else
			{
				const c4<double> c = ee_coeff(p[0], p[1], p[2], p[3]);
				w[0] = c.a, w[1] = c.b, w[2] = -c.c, w[3] = -c.d;
				ka = v[0], kb = v[2];
			}
			v3<double> diff{ 0, 0, 0 };
			if (vf)
			{
				const v3<double> b{ -w[1], -w[2], -w[3] };
				diff = b.x * (p[0] - p[1]) + b.y * (p[0] - p[2]) + b.z * (p[0] - p[3]);
			}
			else
				diff = (w[0] * p[0] + w[1] * p[1]) - ((-w[2]) * p[2] + (-w[3]) * p[3]);
			const double d = sqrt(dot(diff, diff));
			const double dh = ((double)d_hat[ka] + (double)d_hat[kb]) * 0.5;
			const double th = (double)offset[ka] + (double)offset[kb];
			const double stiff = (double)r.area * kappa;
			double		 e, de;
			if (energy == kEnergyQuadratic)
			{
				const double C = (d - th) - dh;
				e = 0.5 * stiff * C * C;
				de = stiff * C;
			}
			else
			{
				e = stiff * barrier(d - th, dh);
				de = stiff * barrier_d1(d - th, dh);
			}
			acc_b += e;
			if (grad && d > 0.0)
			{
				const v3<double> nn = (1.0 / d) * diff;
				for (int k = 0; k < 4; ++k)
				{
					const v3<double> g = (de * w[k]) * nn;
					atomicAdd(&grad[3 * v[k] + 0], g.x);
					atomicAdd(&grad[3 * v[k] + 1], g.y);
					atomicAdd(&grad[3 * v[k] + 2], g.z);
				}
			}
		}
		if ((terms & kTermFriction) && r.lambda != 0.0f && x_start)
		{
			double w[4];
			record_weights(r, w);
			v3<double> rel{ 0, 0, 0 };
			for (int k = 0; k < 4; ++k)
				rel = rel + w[k] * (p[k] - load_v3<double>(x_start, v[k]));
			const v3<double>   nn{ (double)r.normal[0], (double)r.normal[1], (double)r.normal[2] };
			const FrictionEval f = friction_eval((double)r.lambda, nn, rel, kFrictionEps);
			acc_f += f.energy;
			if (grad)
				for (int k = 0; k < 4; ++k)
				{
					atomicAdd(&grad[3 * v[k] + 0], w[k] * f.grad.x);
					atomicAdd(&grad[3 * v[k] + 1], w[k] * f.grad.y);
					atomicAdd(&grad[3 * v[k] + 2], w[k] * f.grad.z);
				}
		}
	}
	__shared__ double s[2][kEnergyThreads / kWave];
	for (int o = 32; o > 0; o >>= 1)
	{
		acc_b += __shfl_xor(acc_b, o, kWave);
		acc_f += __shfl_xor(acc_f, o, kWave);
	}
	if (__lane_id() == 0)
	{
		s[0][threadIdx.x / kWave] = acc_b;
		s[1][threadIdx.x / kWave] = acc_f;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		double tb = 0.0, tf = 0.0;
		for (int k = 0; k < kEnergyThreads / kWave; ++k)
			tb += s[0][k], tf += s[1][k];
		part[blockIdx.x] = tb;
		part[gridDim.x + blockIdx.x] = tf;
	}
}

// fn_process_collision_pair_friction (narrow_phase.cpp:1455-1521), first Newton
// iteration: mu_lambda = -mu k1 with mu the mean of the coefficients of the
// record's vertices 0 and 2 (get_friction_mu, :357-363), zero relative motion.
__global__ void k_friction_init(ContactRec* __restrict__ pairs, const uint32_t* __restrict__ count, uint32_t cap,
	const float* __restrict__ mu)
{
	const uint32_t n = min(*count, cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		ContactRec& r = pairs[i];
		const float m = 0.5f * (mu[r.idx[0] & kIndexMask] + mu[r.idx[2] & kIndexMask]);
		r.lambda = -m * r.k1;
		r.dv[0] = r.dv[1] = r.dv[2] = 0.0f;
	}
}

// The same kernel in later iterations: the first iteration's records keep their
// multiplier and weights, their relative motion since the step start is
// sum_k w_k (x_k - x_start_k), and their repulsion is switched off
// (disable_repulsion_part); written to out (which may be the input).
template <typename P>
__global__ void k_friction_lagged(const ContactRec* pairs, const uint32_t* __restrict__ count,
	uint32_t cap, const P* __restrict__ x, const P* __restrict__ x_start, ContactRec* out)
{
	const uint32_t n = min(*count, cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		ContactRec r = pairs[i];
		double	   w[4];
		record_weights(r, w);
		double rel[3] = { 0, 0, 0 };
		for (int k = 0; k < 4; ++k)
		{
			const uint32_t v = r.idx[k] & kIndexMask;
			for (int c = 0; c < 3; ++c)
				rel[c] += w[k] * ((double)x[3 * v + c] - (double)x_start[3 * v + c]);
		}
		r.dv[0] = (float)rel[0], r.dv[1] = (float)rel[1], r.dv[2] = (float)rel[2];
		r.k1 = r.k2 = 0.0f;
		out[i] = r;
	}
}

__global__ void k_sum_parts(uint32_t n, const double* __restrict__ part, double* __restrict__ out)
{
	if (blockIdx.x == 0 && threadIdx.x == 0)
	{
		double t = 0.0;
		for (uint32_t k = 0; k < n; ++k)
			t += part[k];
		*out = t;
	}
}
} // namespace

void contact_energy(Context* ctx, const ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const float* d_hat, const float* offset, double kappa, int energy, double* out, double* scratch)
{
	UT_REQUIRE(pairs && count && x && d_hat && offset && out && scratch, "contact energy: null input");
	constexpr uint32_t kBlocks = 256;
	k_contact_potential<float><<<kBlocks, kEnergyThreads, 0, ctx->stream>>>(pairs, count, cap, x, nullptr, d_hat,
		offset, kappa, energy, kTermBarrier, nullptr, scratch);
	k_sum_parts<<<1, 64, 0, ctx->stream>>>(kBlocks, scratch, out);
	UT_LAUNCH_CHECK();
}

void contact_potential(Context* ctx, const PotentialArgs& a)
{
	UT_REQUIRE(a.pairs && a.count && (a.x || a.x64) && a.out, "contact potential: null input");
	UT_REQUIRE(!(a.terms & kTermBarrier) || (a.d_hat && a.offset), "contact potential: barrier needs d_hat, offset");
	UT_REQUIRE(!(a.terms & kTermFriction) || (a.x64 ? a.x_start64 != nullptr : a.x_start != nullptr),
		"contact potential: friction needs the lagged positions of the same precision");
	constexpr uint32_t kBlocks = 256;
	double*			   part = a.out + 2;
	if (a.x64)
		k_contact_potential<double><<<kBlocks, kEnergyThreads, 0, ctx->stream>>>(a.pairs, a.count, a.cap, a.x64,
			a.x_start64, a.d_hat, a.offset, a.kappa, a.energy, a.terms, a.grad, part);
	else
		k_contact_potential<float><<<kBlocks, kEnergyThreads, 0, ctx->stream>>>(a.pairs, a.count, a.cap, a.x,
			a.x_start, a.d_hat, a.offset, a.kappa, a.energy, a.terms, a.grad, part);
	k_sum_parts<<<1, 64, 0, ctx->stream>>>(kBlocks, part, a.out);
	k_sum_parts<<<1, 64, 0, ctx->stream>>>(kBlocks, part + kBlocks, a.out + 1);
	UT_LAUNCH_CHECK();
}

void friction_init(Context* ctx, ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* mu)
{
	UT_REQUIRE(pairs && count && mu, "friction init: null input");
	k_friction_init<<<256, 256, 0, ctx->stream>>>(pairs, count, cap, mu);
	UT_LAUNCH_CHECK();
}

void friction_lagged(Context* ctx, const ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const double* x64, const float* x_start, const double* x_start64, ContactRec* out)
{
	UT_REQUIRE(pairs && count && out, "friction lagged: null input");
	UT_REQUIRE(x64 ? x_start64 != nullptr : (x && x_start), "friction lagged: positions of one precision");
	if (x64)
		k_friction_lagged<double><<<256, 256, 0, ctx->stream>>>(pairs, count, cap, x64, x_start64, out);
	else
		k_friction_lagged<float><<<256, 256, 0, ctx->stream>>>(pairs, count, cap, x, x_start, out);
	UT_LAUNCH_CHECK();
}

namespace
{
__global__ void k_cs_leaf_keys(uint32_t n, const uint2* __restrict__ ids, uint32_t* __restrict__ keys,
	uint32_t* __restrict__ vals)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
	{
		keys[i] = ids[i].y;
		vals[i] = i;
	}
}

__global__ void k_cs_query_keys(uint32_t n, uint32_t nvf, int query_bits, const uint2* __restrict__ ids,
	const uint32_t* __restrict__ perm, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < n)
	{
		const uint32_t i = perm[j];
		keys[j] = ((i >= nvf ? 1u : 0u) << query_bits) | ids[i].x;
		vals[j] = i;
	}
}

__global__ void k_cs_gather(uint32_t n, const uint32_t* __restrict__ perm, const ContactRec* __restrict__ recs,
	const uint2* __restrict__ ids, ContactRec* __restrict__ out, uint2* __restrict__ out_ids)
{
	const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < n)
	{
		out[j] = recs[perm[j]];
		out_ids[j] = ids[perm[j]];
	}
}
} // namespace

void ContactSorter::sort(Context* ctx, ContactRec* recs, uint32_t* ids32, uint32_t n, uint32_t nvf, int query_bits,
	int leaf_bits)
{
	UT_REQUIRE(query_bits >= 1 && query_bits <= 31 && leaf_bits >= 1 && leaf_bits <= 32,
		"contact sort: id bits out of range");
	if (n < 2)
		return;
	const hipStream_t s = ctx->stream;
	for (int k = 0; k < 2; ++k)
	{
		keys[k].reserve(n);
		vals[k].reserve(n);
	}
	hist.reserve(radix_sort_hist_size(n));
	tmp.reserve(n);
	tmp_ids.reserve(2 * (size_t)n);
	const uint2* ids = reinterpret_cast<const uint2*>(ids32);
	// LSD: stable by leaf, then stable by (kind, query)
	k_cs_leaf_keys<<<div_up(n, 256), 256, 0, s>>>(n, ids, keys[0].get(), vals[0].get());
	RadixSortBuffers a;
	a.keys[0] = keys[0].get(), a.keys[1] = keys[1].get();
	a.vals[0] = vals[0].get(), a.vals[1] = vals[1].get();
	a.hist = hist.get();
	radix_sort_pairs(a, n, leaf_bits, s);
	const int r = a.result, o = r ^ 1;
	k_cs_query_keys<<<div_up(n, 256), 256, 0, s>>>(n, nvf, query_bits, ids, a.vals[r], a.keys[o], a.vals[o]);
	RadixSortBuffers b;
	b.keys[0] = a.keys[o], b.keys[1] = a.keys[r];
	b.vals[0] = a.vals[o], b.vals[1] = a.vals[r];
	b.hist = hist.get();
	radix_sort_pairs(b, n, query_bits + 1, s);
	k_cs_gather<<<div_up(n, 256), 256, 0, s>>>(n, b.vals[b.result], recs, ids, tmp.get(),
		reinterpret_cast<uint2*>(tmp_ids.get()));
	UT_LAUNCH_CHECK();
	UT_HIP_CHECK(hipMemcpyAsync(recs, tmp.get(), sizeof(ContactRec) * (size_t)n, hipMemcpyDeviceToDevice, s));
	UT_HIP_CHECK(hipMemcpyAsync(ids32, tmp_ids.get(), 8 * (size_t)n, hipMemcpyDeviceToDevice, s));
}

} // namespace ut
