// Contact filtering around intersection contours (see regions.hpp).
#include "regions.hpp"

#include "scan.hpp"
#include "ut_common.hpp"

namespace ut
{

namespace
{

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
		const uint32_t nvv = hit_side_verts(edges, faces, x.type, side == 0 ? x.src : x.dst, side == 0, v);
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
			vals[slot] = ef_vertex_region(k, er);
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

// duplicate removal per vertex, then the kept counts
__global__ void k_region_dedup(uint32_t nv, const uint32_t* __restrict__ off, uint32_t* __restrict__ vals,
	uint32_t* __restrict__ kept)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v < nv)
		kept[v] = region_dedup(vals + off[v], off[v + 1] - off[v]);
}

__global__ void k_region_compact(uint32_t nv, const uint32_t* __restrict__ off, const uint32_t* __restrict__ vals,
	const uint32_t* __restrict__ off2, uint32_t* __restrict__ vals2)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	const uint32_t b = off[v], n = off2[v + 1] - off2[v], o = off2[v];
	for (uint32_t i = 0; i < n; ++i)
		vals2[o + i] = vals[b + i];
}

// fn_identify_invalid_verts_from_proximity_pairs (intersection_resolver.cpp:2353-2374)
__global__ void k_mark_invalid(uint32_t n, const ContactRec* __restrict__ pairs, const float* __restrict__ x,
	const float* __restrict__ offset, uint32_t* __restrict__ invalid)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const RecordSides r = decode_record(pairs[i]);
	if (record_too_close(r, x, offset))
		for (int k = 0; k < 4; ++k)
			invalid[r.id[k]] = 1u;
}

// fn_filter_invalid_proximity_pairs_from_invalid_verts / _from_region_info (:2376-2417)
__global__ void k_disable_pairs(uint32_t n, ContactRec* __restrict__ pairs, RegionView rv)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const RecordSides r = decode_record(pairs[i]);
	if (pair_filtered(rv, r.a, r.na, r.b, r.nb))
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
	off2.reserve(nv + 1);
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
	// one entry per distinct region: the filters' match loops run over these lists
	// for every proximity and swept pair
	k_region_dedup<<<div_up(nv, 256), 256, 0, s>>>(nv, off.get(), regions.get(), cnt.get());
	exclusive_scan(cnt.get(), off2.get(), nv + 1, scan_tmp.get(), nullptr, s);
	regions2.reserve(std::max(total, 1u));
	k_region_compact<<<div_up(nv, 256), 256, 0, s>>>(nv, off.get(), regions.get(), off2.get(), regions2.get());
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
		v.off = off2.get();
		v.regions = regions2.get();
		v.invalid = invalid.get();
	}
	return v;
}

} // namespace ut
