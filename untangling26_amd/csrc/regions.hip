// Contact filtering around intersection contours (see regions.hpp).
#include "regions.hpp"

#include "scan.hpp"
#include "ut_common.hpp"

namespace ut
{

namespace
{

__device__ __forceinline__ uint32_t hit_verts(const uint32_t* edges, const uint32_t* faces, uint32_t type,
	uint32_t prim, bool src, uint32_t* o)
{
	// VF: source vertex / target face; EE: edges; FV: source face / target vertex
	const bool vert = (type == 0 && src) || (type == 2 && !src);
	if (vert)
		return o[0] = prim, 1u;
	if (type == 1)
		return o[0] = edges[2 * prim], o[1] = edges[2 * prim + 1], 2u;
	o[0] = faces[3 * prim], o[1] = faces[3 * prim + 1], o[2] = faces[3 * prim + 2];
	return 3u;
}

// pass 0 counts, pass 1 writes (cursor = running offsets)
__global__ void k_region_hits(int pass, uint32_t n, const PrpHit* __restrict__ h, const uint32_t* __restrict__ hc,
	const uint8_t* __restrict__ suppressed, const uint32_t* __restrict__ edges, const uint32_t* __restrict__ faces,
	uint32_t* __restrict__ cnt, uint32_t* __restrict__ vals)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || suppressed[hc[i]])
		return;
	const PrpHit   x = h[i];
	const uint32_t r = 2 * hc[i];
	uint32_t	   v[3];
	for (int side = 0; side < 2; ++side)
	{
		const uint32_t nvv = hit_verts(edges, faces, x.type, side == 0 ? x.src : x.dst, side == 0, v);
		for (uint32_t k = 0; k < nvv; ++k)
		{
			const uint32_t slot = atomicAdd(&cnt[v[k]], 1u);
			if (pass)
				vals[slot] = r + (uint32_t)side;
		}
	}
}

__global__ void k_region_ef(int pass, uint32_t n, const EfPairRec* __restrict__ pairs,
	const uint32_t* __restrict__ contour_index, const uint32_t* __restrict__ mesh_index, uint32_t* __restrict__ cnt,
	uint32_t* __restrict__ vals)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint32_t er = 2 * contour_index[i] + mesh_index[i];
	for (int k = 0; k < 5; ++k)
	{
		const uint32_t v = pairs[i].idx[k];
		const uint32_t slot = atomicAdd(&cnt[v], 1u);
		if (pass)
			vals[slot] = k < 2 ? er : er ^ 1u;
	}
}

__global__ void k_region_zero(uint32_t n, uint32_t* p)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0u;
}

__global__ void k_region_copy(uint32_t n, const uint32_t* a, uint32_t* b)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		b[i] = a[i];
}

// Record indices carry the collision type in bit 31 (VF: idx[2], EE: idx[3];
// collision_data.h:36-80): every index is masked before it addresses a vertex.
constexpr uint32_t kTypeBit = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7FFFFFFFu;

__device__ __forceinline__ uint32_t rec_sides(const ContactRec& r, uint32_t* a, int* na, uint32_t* b, int* nb)
{
	const bool	   vf = (r.idx[2] & kTypeBit) != 0u;
	const uint32_t i0 = r.idx[0] & kIndexMask, i1 = r.idx[1] & kIndexMask, i2 = r.idx[2] & kIndexMask,
				   i3 = r.idx[3] & kIndexMask;
	if (vf)
	{
		a[0] = i0, *na = 1;
		b[0] = i1, b[1] = i2, b[2] = i3, *nb = 3;
	}
	else
	{
		a[0] = i0, a[1] = i1, *na = 2;
		b[0] = i2, b[1] = i3, *nb = 2;
	}
	return vf ? 1u : 0u;
}

// fn_identify_invalid_verts_from_proximity_pairs (intersection_resolver.cpp:2353-2374)
__global__ void k_mark_invalid(uint32_t n, const ContactRec* __restrict__ pairs, const float* __restrict__ x,
	const float* __restrict__ offset, uint32_t* __restrict__ invalid)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const ContactRec r = pairs[i];
	const bool		 vf = (r.idx[2] & kTypeBit) != 0u;
	const uint32_t	 id[4] = { r.idx[0] & kIndexMask, r.idx[1] & kIndexMask, r.idx[2] & kIndexMask,
		  r.idx[3] & kIndexMask };
	const float		 b0 = r.bary[0], b1 = r.bary[1];
	const float		 w[4] = { vf ? 1.0f : b0, vf ? -b0 : 1.0f - b0, -b1, vf ? b0 + b1 - 1.0f : b1 - 1.0f };
	f3				 d = mk3(0.0f, 0.0f, 0.0f);
	for (int k = 0; k < 4; ++k)
		d = d + ld3(x, id[k]) * w[k];
	const float t = offset[id[0]] + offset[id[2]];
	if (dot3(d, d) < t * t + 1e-10f)
		for (int k = 0; k < 4; ++k)
			invalid[id[k]] = 1u;
}

// fn_filter_invalid_proximity_pairs_from_invalid_verts / _from_region_info (:2376-2417)
__global__ void k_disable_pairs(uint32_t n, ContactRec* __restrict__ pairs, RegionView rv)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint32_t a[3], b[3];
	int		 na, nb;
	rec_sides(pairs[i], a, &na, b, &nb);
	if (pair_filtered(rv, a, na, b, nb))
	{
		pairs[i].k1 = 0.0f;
		pairs[i].k2 = 0.0f;
	}
}

} // namespace

void RegionFilter::init(Context* c, uint32_t nv_)
{
	ctx = c;
	nv = nv_;
	off.reserve(nv + 1);
	cnt.reserve(nv + 1);
	invalid.reserve(nv);
	scan_tmp.reserve(scan_tmp_size(nv + 1));
}

void RegionFilter::build(const PrpResolver& prp, const ContourBuilder& ct)
{
	const hipStream_t s = ctx->stream;
	const uint32_t	  nh = prp.last_final_hits, ne = ct.num_total;
	active = false;
	if (ne == 0)
		return;
	k_region_zero<<<div_up(nv + 1, 256), 256, 0, s>>>(nv + 1, cnt.get());
	if (nh)
		k_region_hits<<<div_up(nh, 256), 256, 0, s>>>(0, nh, prp.final_hits.get(), prp.final_contour.get(),
			prp.suppressed.get(), prp.edges, prp.faces, cnt.get(), nullptr);
	k_region_ef<<<div_up(ne, 256), 256, 0, s>>>(0, ne, ct.pairs.get(), ct.contour_index.get(), ct.mesh_index.get(),
		cnt.get(), nullptr);
	exclusive_scan(cnt.get(), off.get(), nv + 1, scan_tmp.get(), nullptr, s);
	uint32_t total = 0;
	UT_HIP_CHECK(hipMemcpyAsync(&total, off.get() + nv, 4, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	regions.reserve(std::max(total, 1u));
	k_region_copy<<<div_up(nv + 1, 256), 256, 0, s>>>(nv + 1, off.get(), cnt.get());
	if (nh)
		k_region_hits<<<div_up(nh, 256), 256, 0, s>>>(1, nh, prp.final_hits.get(), prp.final_contour.get(),
			prp.suppressed.get(), prp.edges, prp.faces, cnt.get(), regions.get());
	k_region_ef<<<div_up(ne, 256), 256, 0, s>>>(1, ne, ct.pairs.get(), ct.contour_index.get(), ct.mesh_index.get(),
		cnt.get(), regions.get());
	UT_LAUNCH_CHECK();
	active = true;
}

void RegionFilter::filter_proximity(ContactRec* pairs, uint32_t n, const float* x, const float* offset)
{
	const hipStream_t s = ctx->stream;
	k_region_zero<<<div_up(nv, 256), 256, 0, s>>>(nv, invalid.get());
	if (n)
	{
		k_mark_invalid<<<div_up(n, 256), 256, 0, s>>>(n, pairs, x, offset, invalid.get());
		k_disable_pairs<<<div_up(n, 256), 256, 0, s>>>(n, pairs, view());
	}
	UT_LAUNCH_CHECK();
}

RegionView RegionFilter::view() const
{
	RegionView v;
	if (active)
	{
		v.off = off.get();
		v.regions = regions.get();
		v.invalid = invalid.get();
	}
	return v;
}

} // namespace ut
