// extern "C" boundary of libuntangle_hip.  Each entry point converts C++
// exceptions into status codes and records the message on the context.
#include "../../include/untangle.h"
#include "graph_static.hpp"
#include "ut_common.hpp"
#include "ut_runtime.hpp"
#include "lbvh.hpp"
#include "ef_dcd.hpp"
#include "contour.hpp"
#include "pcg.hpp"
#include "contact.hpp"
#include "solver.hpp"
#include "prp.hpp"
#include "abd.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <vector>

// A context is reference counted: the caller's handle plus one reference per
// object created from it, so objects may be destroyed after the context in any
// order (as garbage collectors do at interpreter exit).
struct ut_ctx
{
	ut::Context		 c;
	std::atomic<int> refs{ 1 };
};

namespace
{
void ctx_retain(ut_ctx* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }
void ctx_release(ut_ctx* c)
{
	if (!c || c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1)
		return;
	if (c->c.own_stream && c->c.stream)
	{
		(void)hipSetDevice(c->c.device);
		(void)hipStreamSynchronize(c->c.stream);
		(void)hipStreamDestroy(c->c.stream);
	}
	delete c;
}
} // namespace

struct ut_lbvh
{
	ut::Lbvh b;
	ut_ctx*	 owner = nullptr;
};
struct ut_contours
{
	ut::ContourBuilder b;
	ut_ctx*			   owner = nullptr;
};
struct ut_pcg
{
	ut::PcgSolver s;
	ut_ctx*		  owner = nullptr;
};
struct ut_solver
{
	ut::NewtonSolver s;
	ut_ctx*			 owner = nullptr;
};
struct ut_prp
{
	ut::PrpResolver r;
	ut_ctx*			owner = nullptr;
};
struct ut_abd_pcg
{
	ut::AbdPcg s;
	ut_ctx*	   owner = nullptr;
};

namespace
{
// finish the object's queued work, free it, then drop its context reference
template <typename H>
void destroy_handle(H* h, ut::Context* c)
{
	if (c && c->stream)
	{
		(void)hipSetDevice(c->device);
		(void)hipStreamSynchronize(c->stream);
	}
	ut_ctx* owner = h->owner;
	delete h;
	ctx_release(owner);
}
} // namespace

namespace
{
template <typename F>
int guarded(ut::Context* c, F&& f)
{
	try
	{
		if (c)
			c->activate();
		f();
		return ut::kOk;
	}
	catch (const ut::Error& e)
	{
		if (c)
			c->last_error = e.what();
		return e.code;
	}
	catch (const std::exception& e)
	{
		if (c)
			c->last_error = e.what();
		return ut::kInternal;
	}
	catch (...)
	{
		if (c)
			c->last_error = "unknown exception";
		return ut::kInternal;
	}
}

__global__ void k_morton_points(const float* __restrict__ p, uint32_t n, uint32_t* __restrict__ out)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	auto expand = [](uint32_t v)
	{
		v = (v | (v << 16)) & 0x030000FFu;
		v = (v | (v << 8)) & 0x0300F00Fu;
		v = (v | (v << 4)) & 0x030C30C3u;
		return (v | (v << 2)) & 0x09249249u;
	};
	const float x = fmaxf(fminf(p[3 * i + 0] * 1024.0f, 1023.0f), 0.0f);
	const float y = fmaxf(fminf(p[3 * i + 1] * 1024.0f, 1023.0f), 0.0f);
	const float z = fmaxf(fminf(p[3 * i + 2] * 1024.0f, 1023.0f), 0.0f);
	out[i] = (expand((uint32_t)x) << 2) | (expand((uint32_t)y) << 1) | expand((uint32_t)z);
}

__global__ void k_split_keys(const uint64_t* __restrict__ k, uint32_t n, uint32_t* __restrict__ hi, uint32_t* __restrict__ lo)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	hi[i] = (uint32_t)(k[i] >> 32);
	lo[i] = (uint32_t)(k[i] & 0xFFFFFFFFull);
}
} // namespace

extern "C" {

UT_API int ut_version(void) { return 1; }

UT_API int ut_ctx_create(int device, void* hip_stream, int flags, ut_ctx** out)
{
	if (!out)
		return ut::kInvalid;
	*out = nullptr;
	ut_ctx* h = new (std::nothrow) ut_ctx();
	if (!h)
		return ut::kInternal;
	int rc = guarded(nullptr,
		[&]
		{
			int n = 0;
			UT_HIP_CHECK(hipGetDeviceCount(&n));
			UT_REQUIRE(device >= 0 && device < n, "no such HIP device");
			h->c.device = device;
			UT_HIP_CHECK(hipSetDevice(device));
			ut::graph_static_setup(device);
			if (flags & UT_CTX_OWN_STREAM)
			{
				UT_HIP_CHECK(hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking));
				h->c.own_stream = true;
			}
			else
				h->c.stream = reinterpret_cast<hipStream_t>(hip_stream);
		});
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	*out = h;
	return ut::kOk;
}

UT_API int ut_ctx_destroy(ut_ctx* ctx)
{
	ctx_release(ctx);
	return ut::kOk;
}

UT_API const char* ut_last_error(const ut_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

UT_API int ut_ctx_synchronize(ut_ctx* ctx)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c, [&] { UT_HIP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}

UT_API void* ut_ctx_stream(ut_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

// ---- LBVH ----------------------------------------------------------------
UT_API int ut_lbvh_create(ut_ctx* ctx, int tree_type, uint32_t num_leaves, ut_lbvh** out)
{
	if (!ctx || !out)
		return ut::kInvalid;
	*out = nullptr;
	ut_lbvh* h = new (std::nothrow) ut_lbvh();
	if (!h)
		return ut::kInternal;
	int rc = guarded(&ctx->c, [&] { h->b.init(&ctx->c, tree_type, num_leaves); });
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	h->owner = ctx;
	ctx_retain(ctx);
	*out = h;
	return ut::kOk;
}

UT_API int ut_lbvh_destroy(ut_lbvh* bvh)
{
	if (bvh)
		destroy_handle(bvh, bvh->b.ctx);
	return ut::kOk;
}

UT_API int ut_lbvh_reduce_aabb(ut_lbvh* bvh, const float* x, const uint32_t* prims)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx, [&] { bvh->b.reduce_aabb(x, prims); });
}

UT_API int ut_lbvh_construct_tree(ut_lbvh* bvh)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx, [&] { bvh->b.construct_tree(); });
}

UT_API int ut_lbvh_build_from_sorted_keys(ut_lbvh* bvh, const uint64_t* keys)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx,
		[&]
		{
			UT_REQUIRE(keys, "null keys");
			auto&		   b = bvh->b;
			const uint32_t n = b.num_leaves;
			b.sort.result = 0;
			k_split_keys<<<ut::div_up(n, 256), 256, 0, b.ctx->stream>>>(keys, n, b.sort.keys[0], b.sort.vals[0]);
			UT_LAUNCH_CHECK();
			b.construct_tree_presorted();
		});
}

UT_API int ut_lbvh_refit(ut_lbvh* bvh, const float* thickness, const float* xs, const float* xe, const uint32_t* prims)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx, [&] { bvh->b.update_leaves_and_refit(thickness, xs, xe, prims); });
}

UT_API int ut_lbvh_query(ut_lbvh* bvh, int kind, const float* xb, const float* xe, const uint32_t* edges, uint32_t nq,
	const float* d_hat, const float* thickness, uint32_t* count, uint32_t* list, uint32_t max_count)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx,
		[&] { bvh->b.query(kind, xb, xe, edges, nq, d_hat, thickness, count, list, max_count); });
}

UT_API int ut_lbvh_tree_arrays(ut_lbvh* bvh, uint64_t* morton_sorted, uint32_t* sorted_get_original, uint32_t* parent,
	uint32_t* children, uint32_t* object_idx)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx,
		[&]
		{
			auto&		   b = bvh->b;
			const size_t   n = b.num_leaves, nn = 2 * n - 1;
			hipStream_t	   s = b.ctx->stream;
			const auto	   d2d = hipMemcpyDeviceToDevice;
			if (morton_sorted)
				UT_HIP_CHECK(hipMemcpyAsync(morton_sorted, b.morton_sorted.get(), n * 8, d2d, s));
			if (sorted_get_original)
				UT_HIP_CHECK(hipMemcpyAsync(sorted_get_original, b.sorted_get_original.get(), n * 4, d2d, s));
			if (parent)
				UT_HIP_CHECK(hipMemcpyAsync(parent, b.parent.get(), nn * 4, d2d, s));
			if (children)
				UT_HIP_CHECK(hipMemcpyAsync(children, b.children.get(), nn * 8, d2d, s));
			if (object_idx)
				UT_HIP_CHECK(hipMemcpyAsync(object_idx, b.object_idx.get(), nn * 4, d2d, s));
		});
}

UT_API int ut_lbvh_boxes(ut_lbvh* bvh, float* root_box, float* global_box)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx,
		[&]
		{
			auto&		b = bvh->b;
			hipStream_t s = b.ctx->stream;
			if (root_box)
				UT_HIP_CHECK(hipMemcpyAsync(root_box, b.root_box.get(), 24, hipMemcpyDeviceToDevice, s));
			if (global_box)
				UT_HIP_CHECK(hipMemcpyAsync(global_box, b.global_box.get(), 24, hipMemcpyDeviceToDevice, s));
		});
}

UT_API int ut_lbvh_status(ut_lbvh* bvh, uint32_t* healthy, uint32_t* truncated)
{
	if (!bvh)
		return ut::kInvalid;
	return guarded(bvh->b.ctx,
		[&]
		{
			uint32_t st[2] = { 0, 0 };
			UT_HIP_CHECK(hipMemcpyAsync(st, bvh->b.status.get(), 8, hipMemcpyDeviceToHost, bvh->b.ctx->stream));
			UT_HIP_CHECK(hipStreamSynchronize(bvh->b.ctx->stream));
			if (healthy)
				*healthy = st[0];
			if (truncated)
				*truncated = st[1];
		});
}

UT_API int ut_morton30(ut_ctx* ctx, const float* points, uint32_t n, uint32_t* codes)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(points && codes, "null pointer");
			if (n == 0)
				return;
			k_morton_points<<<ut::div_up(n, 256), 256, 0, ctx->c.stream>>>(points, n, codes);
			UT_LAUNCH_CHECK();
		});
}

// ---- EF DCD ----------------------------------------------------------------
UT_API int ut_ef_dcd(ut_ctx* ctx, const uint32_t* broad, const uint32_t* broad_count, uint32_t broad_cap,
	const float* x, const uint32_t* edges, const uint32_t* faces, const uint32_t* vprop, const float* d_hat,
	uint32_t* count, void* pairs, uint32_t* pair_ids, uint32_t max_pairs)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::EfDcdArgs a;
			a.broad = broad;
			a.broad_count = broad_count;
			a.broad_cap = broad_cap;
			a.x = x;
			a.edges = edges;
			a.faces = faces;
			a.vprop = vprop;
			a.d_hat = d_hat;
			a.count = count;
			a.pairs = reinterpret_cast<uint32_t*>(pairs);
			a.pair_ids = pair_ids;
			a.max_pairs = max_pairs;
			ut::ef_dcd(&ctx->c, a);
		});
}

UT_API int ut_ef_detect(ut_lbvh* face_tree, const float* x, const uint32_t* edges, uint32_t num_edges,
	const uint32_t* faces, const uint32_t* vprop, const float* d_hat, const float* thickness, uint32_t* count,
	void* pairs, uint32_t* pair_ids, uint32_t max_pairs)
{
	if (!face_tree)
		return ut::kInvalid;
	return guarded(face_tree->b.ctx,
		[&]
		{
			ut::EfDcdArgs a;
			a.x = x;
			a.edges = edges;
			a.faces = faces;
			a.vprop = vprop;
			a.d_hat = d_hat;
			a.count = count;
			a.pairs = reinterpret_cast<uint32_t*>(pairs);
			a.pair_ids = pair_ids;
			a.max_pairs = max_pairs;
			ut::ef_detect(face_tree->b.ctx, face_tree->b, a, num_edges, thickness);
		});
}

UT_API int ut_ef_from_proximity(ut_ctx* ctx, const void* recs, const uint32_t* rec_ids, const uint32_t* rec_count,
	uint32_t rec_cap, const float* x, const uint32_t* edges, const uint32_t* faces, const uint32_t* vert_adj_off,
	const uint32_t* vert_adj_edges, const uint32_t* edge_adj_faces, uint32_t* count, void* pairs,
	uint32_t* pair_ids, uint32_t max_pairs)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::EfFromProximityArgs a;
			a.recs = reinterpret_cast<const uint32_t*>(recs);
			a.rec_ids = rec_ids;
			a.rec_count = rec_count;
			a.rec_cap = rec_cap;
			a.x = x;
			a.edges = edges;
			a.faces = faces;
			a.vert_adj_off = vert_adj_off;
			a.vert_adj_edges = vert_adj_edges;
			a.edge_adj_faces = edge_adj_faces;
			a.count = count;
			a.pairs = reinterpret_cast<uint32_t*>(pairs);
			a.pair_ids = pair_ids;
			a.max_pairs = max_pairs;
			ut::ef_from_proximity(&ctx->c, a);
		});
}

// ---- contours ---------------------------------------------------------------
UT_API int ut_contours_create(ut_ctx* ctx, uint32_t nv, uint32_t ne, uint32_t nf, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, ut_contours** out)
{
	if (!ctx || !out)
		return ut::kInvalid;
	*out = nullptr;
	ut_contours* h = new (std::nothrow) ut_contours();
	if (!h)
		return ut::kInternal;
	int rc = guarded(&ctx->c, [&] { h->b.init(&ctx->c, nv, ne, nf, edges, faces, face_adj_edges, edge_adj_faces); });
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	h->owner = ctx;
	ctx_retain(ctx);
	*out = h;
	return ut::kOk;
}

UT_API int ut_contours_destroy(ut_contours* ct)
{
	if (ct)
		destroy_handle(ct, ct->b.ctx);
	return ut::kOk;
}

UT_API int ut_contours_build(ut_contours* ct, uint32_t num_pairs, const void* pairs, const uint32_t* pair_ids,
	int ignore_near_zero_dist_pairs, uint32_t* num_total, uint32_t* num_contours)
{
	if (!ct)
		return ut::kInvalid;
	return guarded(ct->b.ctx,
		[&]
		{
			ct->b.build(num_pairs, reinterpret_cast<const ut::EfPairRec*>(pairs), pair_ids,
				ignore_near_zero_dist_pairs == 0);
			if (num_total)
				*num_total = ct->b.num_total;
			if (num_contours)
				*num_contours = ct->b.num_contours;
		});
}

// One UT_CT_* output of a contour build: device source and byte count.
static void contour_output(const ut::ContourBuilder& b, int what, const void** src, size_t* bytes)
{
	const size_t n = b.num_total;
	switch (what)
	{
	case UT_CT_PAIR_IDS: *src = b.ids.get(), *bytes = 8 * n; break;
	case UT_CT_PAIRS: *src = b.pairs.get(), *bytes = sizeof(ut::EfPairRec) * n; break;
	case UT_CT_CONTOUR_INDEX: *src = b.contour_index.get(), *bytes = 4 * n; break;
	case UT_CT_MESH_INDEX: *src = b.mesh_index.get(), *bytes = 4 * n; break;
	case UT_CT_ADJ: *src = b.adj.get(), *bytes = 4 * (size_t)b.num_adj; break;
	case UT_CT_ADJ_N: *src = b.adj_n.get(), *bytes = 4 * n; break;
	case UT_CT_INCONSISTENT: *src = b.comp_inconsistent.get(), *bytes = 4 * n; break;
	case UT_CT_COMP_MIN: *src = b.comp_min.get(), *bytes = 4 * n; break;
	case UT_CT_ADJ_OFF: *src = b.adj_off.get(), *bytes = n ? 4 * (n + 1) : 0; break;
	case UT_CT_CONTOUR_TYPE: *src = b.contour_type.get(), *bytes = 4 * (size_t)b.num_contours; break;
	case UT_CT_MESH_CLOSED: *src = b.mesh_closed.get(), *bytes = 8 * (size_t)b.num_contours; break;
	default: throw ut::Error(ut::kInvalid, "unknown contour output");
	}
}

static void copy_contour_output(const ut::ContourBuilder& b, int what, void* dst)
{
	const void* src = nullptr;
	size_t		bytes = 0;
	contour_output(b, what, &src, &bytes);
	UT_REQUIRE(dst || bytes == 0, "null destination");
	if (bytes)
		UT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b.ctx->stream));
}

UT_API int ut_contours_get(ut_contours* ct, int what, void* dst)
{
	if (!ct)
		return ut::kInvalid;
	return guarded(ct->b.ctx, [&] { copy_contour_output(ct->b, what, dst); });
}

UT_API int ut_contours_counts(ut_contours* ct, uint32_t* out4)
{
	if (!ct || !out4)
		return ut::kInvalid;
	out4[0] = ct->b.num_input;
	out4[1] = ct->b.num_pairs;
	out4[2] = ct->b.num_total;
	out4[3] = ct->b.num_adj;
	return ut::kOk;
}

UT_API int ut_graph_components(ut_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
	const uint32_t* flip_w, uint32_t* scratch, uint32_t* comp_min, uint32_t* parity, uint32_t* inconsistent)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(n == 0 || (row_ptr && scratch && comp_min && parity && inconsistent), "null pointer");
			ut::graph_components(&ctx->c, n, row_ptr, col_idx, flip_w, scratch, comp_min, parity, inconsistent, nullptr);
		});
}

UT_API int ut_graph_components_static(ut_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
	const uint32_t* flip_w, uint32_t* comp_min, uint32_t* parity, uint32_t* inconsistent)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(n <= ut::kGraphStaticMaxVertices,
				"ut_graph_components_static: n = " + std::to_string(n) + " exceeds the static scratch limit of "
					+ std::to_string(ut::kGraphStaticMaxVertices) + " vertices");
			UT_REQUIRE(n == 0 || (row_ptr && comp_min && parity && inconsistent), "null pointer");
			uint32_t* const uf = ut::graph_static_scratch(ctx->c.device);
			ut::graph_components(&ctx->c, n, row_ptr, col_idx, flip_w, uf, comp_min, parity, inconsistent, nullptr);
		});
}

// ---- PCG ---------------------------------------------------------------------
UT_API int ut_pcg_create(ut_ctx* ctx, uint32_t n, uint32_t nslices, uint32_t nslots, const uint32_t* slice_off,
	const uint32_t* col, const uint32_t* diag_slot, ut_pcg** out)
{
	if (!ctx || !out)
		return ut::kInvalid;
	*out = nullptr;
	ut_pcg* h = new (std::nothrow) ut_pcg();
	if (!h)
		return ut::kInternal;
	int rc = guarded(&ctx->c, [&] { h->s.init(&ctx->c, n, nslices, nslots, slice_off, col, diag_slot); });
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	h->owner = ctx;
	ctx_retain(ctx);
	*out = h;
	return ut::kOk;
}

UT_API int ut_pcg_destroy(ut_pcg* pcg)
{
	if (pcg)
		destroy_handle(pcg, pcg->s.ctx);
	return ut::kOk;
}

static int pcg_solve_impl(ut_pcg* pcg, const ut_pcg_params* params, const void* vals, const double* vals64,
	const double* b, double* x, ut_pcg_report* report)
{
	if (!pcg || !params)
		return ut::kInvalid;
	ut::PcgReport rep;
	int			  rc = guarded(pcg->s.ctx,
		 [&]
		 {
			 ut::PcgParams p;
			 p.precision = params->precision;
			 p.preconditioner = params->preconditioner;
			 p.max_iters = params->max_iters;
			 p.rel_tol = params->rel_tol;
			 p.inner_tol = params->inner_tol;
			 p.max_refine = params->max_refine;
			 p.regularization = params->regularization;
			 p.check_every = params->check_every ? params->check_every : 8;
			 UT_REQUIRE(params->recurrence == ut::kPcgClassic || params->recurrence == ut::kPcgSingleReduction,
				 "pcg: unknown recurrence");
			 p.recurrence = params->recurrence;
			 rep = pcg->s.solve(p, vals, b, x, vals64);
		 });
	if (rc != ut::kOk)
		return rc;
	if (report)
	{
		report->relative_residual = rep.relative_residual;
		report->infinity_norm = rep.infinity_norm;
		report->iterations = rep.iterations;
		report->refinements = rep.refinements;
		report->converged = rep.converged;
		report->failure = rep.failure;
	}
	return rep.converged ? ut::kOk : ut::kNotConverged;
}

UT_API int ut_pcg_solve(ut_pcg* pcg, const ut_pcg_params* params, const void* vals, const double* b, double* x,
	ut_pcg_report* report)
{
	return pcg_solve_impl(pcg, params, vals, nullptr, b, x, report);
}

UT_API int ut_pcg_solve_refined(ut_pcg* pcg, const ut_pcg_params* params, const float* vals32, const double* vals64,
	const double* b, double* x, ut_pcg_report* report)
{
	if (!vals64)
		return ut::kInvalid;
	return pcg_solve_impl(pcg, params, vals32, vals64, b, x, report);
}

UT_API int ut_bsr_spmv(ut_pcg* pcg, const void* vals, int vals_fp32, const double* x, double* y)
{
	if (!pcg)
		return ut::kInvalid;
	return guarded(pcg->s.ctx, [&] { pcg->s.spmv64(vals, vals_fp32 != 0, x, y); });
}

UT_API int ut_pcg_set_contacts(ut_pcg* pcg, const void* pairs, const uint32_t* count, uint32_t cap)
{
	if (!pcg || (pairs && !count))
		return ut::kInvalid;
	pcg->s.contacts = reinterpret_cast<const ut::ContactRec*>(pairs);
	pcg->s.contact_count = pairs ? count : nullptr;
	pcg->s.contact_cap = pairs ? cap : 0u;
	return ut::kOk;
}

// ---- contacts ----------------------------------------------------------------
UT_API int ut_contact_energy(ut_ctx* ctx, const void* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const float* d_hat, const float* offset, double kappa, int energy, double* out)
{
	if (!ctx || !pairs || !count || !x || !d_hat || !offset || !out)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(energy == ut::kEnergyQuadratic || energy == ut::kEnergyBarrier, "contact energy: unknown energy");
			ut::contact_energy(&ctx->c, reinterpret_cast<const ut::ContactRec*>(pairs), count, cap, x, d_hat, offset,
				kappa, energy, out, out + 1);
		});
}

UT_API int ut_friction_init(ut_ctx* ctx, void* pairs, const uint32_t* count, uint32_t cap, const float* friction_mu)
{
	if (!ctx || !pairs || !count || !friction_mu)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&] { ut::friction_init(&ctx->c, reinterpret_cast<ut::ContactRec*>(pairs), count, cap, friction_mu); });
}

UT_API int ut_friction_lagged(ut_ctx* ctx, const void* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const double* x64, const float* x_start, const double* x_start64, void* out)
{
	if (!ctx || !pairs || !count || !out || (x64 ? !x_start64 : (!x || !x_start)))
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::friction_lagged(&ctx->c, reinterpret_cast<const ut::ContactRec*>(pairs), count, cap, x, x64, x_start,
				x_start64, reinterpret_cast<ut::ContactRec*>(out));
		});
}

UT_API int ut_contact_potential(ut_ctx* ctx, const ut_potential_args* q)
{
	if (!ctx || !q || !q->pairs || !q->count || !q->out || (!q->x && !q->x64))
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(q->energy == ut::kEnergyQuadratic || q->energy == ut::kEnergyBarrier,
				"contact potential: unknown energy");
			UT_REQUIRE(q->terms > 0 && q->terms <= 3, "contact potential: terms must be UT_TERM_BARRIER and/or "
													  "UT_TERM_FRICTION");
			ut::PotentialArgs a;
			a.pairs = reinterpret_cast<const ut::ContactRec*>(q->pairs);
			a.count = q->count;
			a.cap = q->cap;
			a.x = q->x64 ? nullptr : q->x;
			a.x64 = q->x64;
			a.x_start = q->x_start;
			a.x_start64 = q->x_start64;
			a.d_hat = q->d_hat;
			a.offset = q->offset;
			a.kappa = q->kappa;
			a.energy = q->energy;
			a.terms = q->terms;
			a.grad = q->grad;
			a.out = q->out;
			ut::contact_potential(&ctx->c, a);
		});
}

// ---- affine bodies -------------------------------------------------------------
namespace
{
ut::AbdDesc abd_desc(const ut_abd_desc* q)
{
	UT_REQUIRE(q && q->nb && q->nv && q->body && q->xbar && q->mass4 && q->volume && q->kappa,
		"abd: incomplete body description");
	ut::AbdDesc d;
	d.nb = q->nb, d.nv = q->nv, d.vert_off = q->vert_off, d.body = q->body, d.xbar = q->xbar;
	d.mass4 = q->mass4, d.volume = q->volume, d.kappa = q->kappa;
	return d;
}
} // namespace

UT_API int ut_abd_q_to_x(ut_ctx* ctx, const ut_abd_desc* desc, const double* q, float* x32, double* x64)
{
	if (!ctx || !desc || !q || (!x32 && !x64))
		return ut::kInvalid;
	return guarded(&ctx->c, [&] { ut::abd_q_to_x(&ctx->c, abd_desc(desc), q, x32, x64); });
}

UT_API int ut_abd_assemble(ut_ctx* ctx, const ut_abd_desc* desc, const ut_abd_assemble_args* q)
{
	if (!ctx || !desc || !q || !q->q || !q->grad)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::AbdAssembleArgs a;
			a.q = q->q;
			a.q_tilde = q->q_tilde;
			a.inv_h2 = q->inv_h2;
			a.inertia_grad_scale = q->quasi_static ? 0.0 : 1.0;
			a.inertia_hess_scale = (q->quasi_static && q->qs_no_inertia_hessian) ? 0.0 : 1.0;
			a.contacts = reinterpret_cast<const ut::ContactRec*>(q->contacts);
			a.contact_count = q->contact_count;
			a.contact_cap = q->contacts ? q->contact_cap : 0u;
			a.grad = q->grad;
			a.blocks = q->blocks;
			a.energy = q->energy;
			ut::abd_assemble(&ctx->c, abd_desc(desc), a);
		});
}

UT_API int ut_abd_gather(ut_ctx* ctx, const ut_abd_desc* desc, const double* gx, double* gq, int accumulate)
{
	if (!ctx || !desc || !gx || !gq || !desc->vert_off)
		return ut::kInvalid;
	return guarded(&ctx->c, [&] { ut::abd_gather(&ctx->c, abd_desc(desc), gx, gq, accumulate != 0); });
}

UT_API int ut_abd_apply(ut_ctx* ctx, const ut_abd_desc* desc, const double* blocks, const void* contacts,
	const uint32_t* contact_count, uint32_t contact_cap, const double* p, double* y)
{
	if (!ctx || !desc || !blocks || !p || !y || (contacts && !contact_count))
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::abd_apply(&ctx->c, abd_desc(desc), blocks, reinterpret_cast<const ut::ContactRec*>(contacts),
				contacts ? contact_count : nullptr, contacts ? contact_cap : 0u, p, y);
		});
}

UT_API int ut_abd_pcg_create(ut_ctx* ctx, uint32_t nb, ut_abd_pcg** out)
{
	if (!ctx || !out || nb == 0)
		return ut::kInvalid;
	*out = nullptr;
	ut_abd_pcg* h = new (std::nothrow) ut_abd_pcg;
	if (!h)
		return ut::kInternal;
	const int rc = guarded(&ctx->c, [&] { h->s.init(&ctx->c, nb); });
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	h->owner = ctx;
	ctx_retain(ctx);
	*out = h;
	return ut::kOk;
}

UT_API int ut_abd_pcg_destroy(ut_abd_pcg* pcg)
{
	if (!pcg)
		return ut::kInvalid;
	destroy_handle(pcg, pcg->s.ctx);
	return ut::kOk;
}

UT_API int ut_abd_pcg_solve(ut_abd_pcg* pcg, const ut_abd_desc* desc, const double* blocks, const void* contacts,
	const uint32_t* contact_count, uint32_t contact_cap, const double* b, double* x, int precision,
	uint32_t max_iters, double rel_tol, ut_pcg_report* report)
{
	if (!pcg || !desc || !blocks || !b || !x || (contacts && !contact_count))
		return ut::kInvalid;
	ut::AbdPcgReport r;
	const int		 rc = guarded(pcg->s.ctx,
				[&]
				{
			UT_REQUIRE(precision == ut::kPcgFp64 || precision == ut::kPcgMixed,
				"abd pcg: precision must be UT_PCG_FP64 or UT_PCG_MIXED");
			r = pcg->s.solve(abd_desc(desc), blocks, reinterpret_cast<const ut::ContactRec*>(contacts),
				contacts ? contact_count : nullptr, contacts ? contact_cap : 0u, b, x, precision, max_iters, rel_tol);
				});
	if (rc != ut::kOk)
		return rc;
	if (report)
	{
		std::memset(report, 0, sizeof(*report));
		report->relative_residual = r.relative_residual;
		report->iterations = r.iterations;
		report->refinements = r.refinements;
		report->converged = r.converged;
		report->failure = r.failure;
	}
	return r.converged ? ut::kOk : ut::kNotConverged;
}

UT_API int ut_contact_query(ut_ctx* ctx, int kind, const ut_contact_args* args, ut_lbvh* tree,
	const uint32_t* broad, const uint32_t* broad_count, uint32_t broad_cap)
{
	if (!ctx || !args)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::ContactArgs a;
			a.num_verts = args->num_verts;
			a.num_edges = args->num_edges;
			a.x = args->x;
			a.x_end = args->x_end;
			a.edges = args->edges;
			a.faces = args->faces;
			a.vprop = args->vprop;
			a.d_hat = args->d_hat;
			a.offset = args->offset;
			a.rest_vert_area = args->rest_vert_area;
			a.rest_edge_area = args->rest_edge_area;
			a.rest_face_area = args->rest_face_area;
			a.kappa = args->kappa;
			a.energy = args->energy;
			a.count = args->count;
			a.pairs = reinterpret_cast<ut::ContactRec*>(args->pairs);
			a.pair_ids = reinterpret_cast<uint2*>(args->pair_ids);
			a.max_pairs = args->max_pairs;
			a.toi = args->toi;
			ut::contact_query(&ctx->c, kind, a, tree ? &tree->b : nullptr, broad, broad_count, broad_cap);
		});
}

UT_API int ut_toi_reset(ut_ctx* ctx, float* toi)
{
	if (!ctx)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(toi, "null toi");
			ut::toi_reset(&ctx->c, toi);
		});
}

// ---- Newton solver -------------------------------------------------------------
namespace
{
ut::PcgParams to_pcg(const ut_pcg_params& q)
{
	ut::PcgParams p;
	p.precision = q.precision;
	p.preconditioner = q.preconditioner;
	p.max_iters = q.max_iters;
	p.rel_tol = q.rel_tol;
	p.inner_tol = q.inner_tol;
	p.max_refine = q.max_refine;
	p.regularization = q.regularization;
	p.check_every = q.check_every ? q.check_every : 8;
	UT_REQUIRE(q.recurrence == ut::kPcgClassic || q.recurrence == ut::kPcgSingleReduction, "pcg: unknown recurrence");
	p.recurrence = q.recurrence;
	return p;
}
ut::SolverParams to_params(const ut_solver_params& q)
{
	ut::SolverParams p;
	p.dt = q.dt;
	for (int k = 0; k < 3; ++k)
		p.gravity[k] = q.gravity[k];
	p.kappa = q.kappa;
	p.energy = q.energy;
	p.damping = q.damping;
	p.nonlinear_iters = q.nonlinear_iters;
	p.use_self_collision = q.use_self_collision;
	p.use_ccd = q.use_ccd;
	p.use_untangling = q.use_untangling;
	p.pcg = to_pcg(q.pcg);
	p.use_prp = q.use_prp;
	p.quasi_static = q.quasi_static;
	p.qs_inertia_hessian = 1;
	p.region_filter = q.region_filter;
	p.ignore_near_zero_dist_pairs = q.ignore_near_zero_dist_pairs;
	if (q.prp.total_dirs)
	{
		p.prp.max_contours = q.prp.max_contours;
		p.prp.total_dirs = q.prp.total_dirs;
		p.prp.batch_contours = q.prp.batch_contours ? q.prp.batch_contours : 32;
		p.prp.response_depth = q.prp.response_depth;
		p.prp.kappa = q.prp.kappa;
		p.prp.tau = q.prp.tau;
	}
	return p;
}
} // namespace

UT_API int ut_solver_create(ut_ctx* ctx, const ut_solver_desc* desc, const ut_solver_params* params, const float* x0,
	const float* v0, ut_solver** out)
{
	if (!ctx || !desc || !params || !out)
		return ut::kInvalid;
	*out = nullptr;
	ut_solver* h = new (std::nothrow) ut_solver();
	if (!h)
		return ut::kInternal;
	int rc = guarded(&ctx->c,
		[&]
		{
			UT_REQUIRE(x0, "null initial positions");
			ut::SolverDesc d;
			d.nv = desc->nv;
			d.ne = desc->ne;
			d.nf = desc->nf;
			d.nh = desc->nh;
			d.edges = desc->edges;
			d.faces = desc->faces;
			d.face_adj_edges = desc->face_adj_edges;
			d.edge_adj_faces = desc->edge_adj_faces;
			d.mass = desc->mass;
			d.d_hat = desc->d_hat;
			d.offset = desc->offset;
			d.vprop = desc->vprop;
			d.rest_vert_area = desc->rest_vert_area;
			d.rest_edge_area = desc->rest_edge_area;
			d.rest_face_area = desc->rest_face_area;
			d.dm = desc->dm;
			d.mu_lambda = desc->mu_lambda;
			d.hinges = desc->hinges;
			d.rest_angle = desc->rest_angle;
			d.bend_stiff_area = desc->bend_stiff_area;
			d.nslices = desc->nslices;
			d.nslots = desc->nslots;
			d.slice_off = desc->slice_off;
			d.col = desc->col;
			d.diag_slot = desc->diag_slot;
			d.face_slots = desc->face_slots;
			d.hinge_slots = desc->hinge_slots;
			d.rest_x = desc->rest_x;
			d.friction_mu = desc->friction_mu;
			if (desc->abd)
			{
				d.abd = abd_desc(desc->abd);
				d.q0 = desc->q0;
			}
			h->s.init(&ctx->c, d, to_params(*params), x0, v0);
		});
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	h->owner = ctx;
	ctx_retain(ctx);
	*out = h;
	return ut::kOk;
}

UT_API int ut_solver_destroy(ut_solver* s)
{
	if (s)
		destroy_handle(s, s->s.ctx);
	return ut::kOk;
}

UT_API int ut_solver_step(ut_solver* s, ut_step_report* report)
{
	if (!s)
		return ut::kInvalid;
	ut::StepReport r;
	int			   rc = guarded(s->s.ctx, [&] { r = s->s.physics_step_GPU(); });
	if (rc != ut::kOk)
		return rc;
	if (report)
	{
		report->vf_pairs = r.vf_pairs;
		report->ee_pairs = r.ee_pairs;
		report->ef_pairs = r.ef_pairs;
		report->ef_total = r.ef_total;
		report->contours = r.contours;
		report->pcg_iterations = r.pcg_iterations;
		report->pcg_refinements = r.pcg_refinements;
		report->pcg_relative_residual = r.pcg_relative_residual;
		report->pcg_converged = r.pcg_converged;
		report->toi = r.toi;
		for (int k = 0; k < 3; ++k)
			report->energy[k] = r.energy[k];
		report->ms_detect = r.ms_detect;
		report->ms_contours = r.ms_contours;
		report->ms_assemble = r.ms_assemble;
		report->ms_pcg = r.ms_pcg;
		report->ms_ccd = r.ms_ccd;
		report->ms_total = r.ms_total;
		report->prp_responses = r.prp_responses;
		report->prp_contours = r.prp_contours;
		report->prp_levels = r.prp_levels;
		report->prp_hits = r.prp_hits;
		report->ms_prp = r.ms_prp;
		report->prp_slots = r.prp_slots;
		report->prp_rays_vf = r.prp_rays[0], report->prp_rays_ee = r.prp_rays[1], report->prp_rays_fv = r.prp_rays[2];
		report->prp_ee_launches = r.prp_ee.launches, report->ms_prp_ee = r.prp_ee.ms;
		report->prp_ee_requests = r.prp_ee.requests, report->prp_ee_visits = r.prp_ee.visits;
		report->prp_ee_leaf_tests = r.prp_ee.leaf_tests, report->prp_ee_hits = r.prp_ee.hits;
		report->prp_ee_cand_tests = r.prp_ee.cand_tests, report->prp_ee_wave_steps = r.prp_ee.wave_steps;
	}
	return ut::kOk;
}

UT_API int ut_solver_set_params(ut_solver* s, const ut_solver_params* params)
{
	if (!s || !params)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			s->s.set_params(to_params(*params));
		});
}

UT_API int ut_solver_get_state(ut_solver* s, float* x, float* v)
{
	if (!s)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			const size_t bytes = 12 * (size_t)s->s.d.nv;
			if (x)
				UT_HIP_CHECK(hipMemcpyAsync(x, s->s.x.get(), bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
			if (v)
				UT_HIP_CHECK(hipMemcpyAsync(v, s->s.v.get(), bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
		});
}

UT_API int ut_solver_set_state(ut_solver* s, const float* x, const float* v)
{
	if (!s)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			const size_t bytes = 12 * (size_t)s->s.d.nv;
			if (x)
				UT_HIP_CHECK(hipMemcpyAsync(s->s.x.get(), x, bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
			if (v)
				UT_HIP_CHECK(hipMemcpyAsync(s->s.v.get(), v, bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
		});
}

UT_API int ut_solver_contour_get(ut_solver* s, int what, void* dst)
{
	if (!s)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			UT_REQUIRE(s->s.contours.ctx, "solver has no contour builder (topology not given)");
			copy_contour_output(s->s.contours, what, dst);
		});
}

UT_API int ut_assemble(ut_ctx* ctx, const ut_assemble_args* q)
{
	if (!ctx || !q)
		return ut::kInvalid;
	return guarded(&ctx->c,
		[&]
		{
			ut::AssembleArgs a;
			a.nv = q->nv;
			a.nf = q->nf;
			a.nh = q->nh;
			a.nslots = q->nslots;
			a.x = q->x;
			a.x64 = q->x64;
			a.x_tilde = q->x_tilde;
			a.mass = q->mass;
			a.inv_h2 = q->inv_h2;
			a.inertia_grad_scale = q->quasi_static ? 0.0 : 1.0;
			a.inertia_hess_scale = (q->quasi_static && q->qs_no_inertia_hessian) ? 0.0 : 1.0;
			a.faces = q->faces;
			a.dm = q->dm;
			a.face_area = q->face_area;
			a.mu_lambda = q->mu_lambda;
			a.face_slots = q->face_slots;
			a.hinges = q->hinges;
			a.rest_angle = q->rest_angle;
			a.bend_stiff_area = q->bend_stiff_area;
			a.hinge_slots = q->hinge_slots;
			a.contacts = reinterpret_cast<const ut::ContactRec*>(q->contacts);
			a.contact_count = q->contact_count;
			a.contact_cap = q->contact_cap;
			a.diag_slot = q->diag_slot;
			a.grad = q->grad;
			a.vals = q->vals;
			a.vals32 = q->vals32;
			a.energy = q->energy;
			ut::assemble(&ctx->c, a);
		});
}

} // extern "C"

/* ------------------------------------------------------------------ PRP */

static_assert(sizeof(ut_prp_slot) == sizeof(ut::PrpSlot), "slot layout");
static_assert(sizeof(ut_prp_summary) == sizeof(ut::PrpSummary), "summary layout");
static_assert(sizeof(ut_prp_hit) == sizeof(ut::PrpHit), "hit layout");

static ut::PrpParams prp_params(const ut_prp_params* p)
{
	ut::PrpParams q;
	if (p)
	{
		q.max_contours = p->max_contours;
		q.total_dirs = p->total_dirs;
		q.batch_contours = p->batch_contours;
		q.response_depth = p->response_depth;
		q.kappa = p->kappa;
		q.tau = p->tau;
	}
	return q;
}

UT_API int ut_prp_create(ut_ctx* ctx, uint32_t nv, uint32_t ne, uint32_t nf, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, const float* rest_x,
	const float* vert_area, const float* edge_area, const float* face_area, const float* mass,
	const ut_prp_params* params, ut_prp** out)
{
	if (!ctx || !out || !edges || !faces || !face_adj_edges || !edge_adj_faces || !rest_x || !vert_area
		|| !edge_area || !face_area || !mass)
		return ut::kInvalid;
	*out = nullptr;
	ut_prp* h = new (std::nothrow) ut_prp();
	if (!h)
		return ut::kInternal;
	int rc = guarded(&ctx->c,
		[&]
		{
			h->r.init(&ctx->c, nv, ne, nf, edges, faces, face_adj_edges, edge_adj_faces, rest_x, vert_area, edge_area,
				face_area, mass, prp_params(params));
		});
	if (rc != ut::kOk)
	{
		delete h;
		return rc;
	}
	h->owner = ctx;
	ctx_retain(ctx);
	*out = h;
	return ut::kOk;
}

UT_API int ut_prp_destroy(ut_prp* p)
{
	if (p)
		destroy_handle(p, p->r.ctx);
	return ut::kOk;
}

UT_API int ut_prp_set_params(ut_prp* p, const ut_prp_params* params)
{
	if (!p || !params)
		return ut::kInvalid;
	return guarded(p->r.ctx,
		[&]
		{
			const ut::PrpParams q = prp_params(params);
			UT_REQUIRE(q.total_dirs >= 1 && q.total_dirs <= 64, "prp: total_dirs must be in [1, 64]");
			UT_REQUIRE(q.batch_contours >= 1, "prp: batch_contours must be positive");
			p->r.params = q;
		});
}

UT_API int ut_prp_resolve(ut_prp* p, const float* x, ut_contours* ct, ut_lbvh* face_tree, ut_lbvh* edge_tree,
	void* out_pairs, uint32_t out_cap, uint32_t* num_pairs, ut_prp_choice* choices, uint32_t choices_cap,
	ut_prp_stats* stats)
{
	if (!p || !x || !ct || !face_tree || !edge_tree || (!out_pairs && out_cap) || !num_pairs)
		return ut::kInvalid;
	uint32_t nresp = 0;
	int		 rc = guarded(p->r.ctx,
		 [&]
		 {
			 std::vector<ut::PrpChoice> ch;
			 ut::PrpStats			   st;
			 nresp = p->r.resolve(x, ct->b, face_tree->b, edge_tree->b, ch, &st);
			 *num_pairs = nresp;
			 const uint32_t n = std::min(nresp, out_cap);
			 if (n)
			 {
				 UT_HIP_CHECK(hipMemcpyAsync(out_pairs, p->r.responses.get(), sizeof(ut::ContactRec) * n,
					 hipMemcpyDeviceToDevice, p->r.ctx->stream));
				 UT_HIP_CHECK(hipStreamSynchronize(p->r.ctx->stream));
			 }
			if (choices)
				for (uint32_t c = 0; c < ch.size() && c < choices_cap; ++c)
				{
					const ut::PrpChoice& a = ch[c];
					ut_prp_choice&		 b = choices[c];
					b.combo = a.combo, b.ks = a.ks, b.kd = a.kd, b.hit_count = a.hit_count, b.covered = a.covered;
					b.ef_count = a.ef_count, b.levels = a.levels, b.suppressed = a.suppressed;
					for (int k = 0; k < 3; ++k)
						b.dir[k] = a.dir[k];
					b.depth = a.depth, b.objective = a.objective, b.area = a.area, b.disp_cost = a.disp_cost;
				}
			if (stats)
			{
				stats->contours = st.contours, stats->levels = st.levels, stats->slots = st.slots;
				stats->hits = st.hits, stats->responses = st.responses;
				stats->ms_prepare = st.ms_prepare, stats->ms_levels = st.ms_levels, stats->ms_final = st.ms_final;
				stats->rays_vf = st.rays[0], stats->rays_ee = st.rays[1], stats->rays_fv = st.rays[2];
			}
		});
	if (rc != ut::kOk)
		return rc;
	return nresp > out_cap ? ut::kCapacity : ut::kOk;
}

UT_API int ut_prp_prepare(ut_prp* p, const float* x, ut_contours* ct, uint32_t c0, uint32_t n)
{
	if (!p || !x || !ct)
		return ut::kInvalid;
	return guarded(p->r.ctx, [&] { p->r.prepare(x, ct->b, c0, n); });
}

UT_API int ut_prp_directions(ut_prp* p, uint32_t contour, float* dirs, uint32_t* count)
{
	if (!p || !count)
		return ut::kInvalid;
	return guarded(p->r.ctx,
		[&]
		{
			UT_REQUIRE(contour < p->r.h_dirs.size(), "prp: contour out of range");
			const auto& d = p->r.h_dirs[contour];
			*count = (uint32_t)(d.size() / 3);
			if (dirs)
				std::copy(d.begin(), d.end(), dirs);
		});
}

UT_API int ut_prp_distances(ut_prp* p, uint32_t local_contour, uint32_t role, uint32_t* hop, float* geo,
	uint32_t* diameter)
{
	if (!p || role > 1)
		return ut::kInvalid;
	return guarded(p->r.ctx,
		[&]
		{
			auto&		   r = p->r;
			UT_REQUIRE(local_contour < r.batch_n, "prp: contour not in the prepared batch");
			const size_t   off = (size_t)(2 * local_contour + role) * r.nv;
			const hipStream_t s = r.ctx->stream;
			if (hop)
				UT_HIP_CHECK(hipMemcpyAsync(hop, r.hop_b + off, 4 * (size_t)r.nv, hipMemcpyDeviceToHost, s));
			if (geo)
				UT_HIP_CHECK(hipMemcpyAsync(geo, r.geo_b + off, 4 * (size_t)r.nv, hipMemcpyDeviceToHost, s));
			UT_HIP_CHECK(hipStreamSynchronize(s));
			if (diameter)
				*diameter = r.diam_b[2 * local_contour + role];
		});
}

UT_API int ut_prp_eval(ut_prp* p, ut_lbvh* face_tree, ut_lbvh* edge_tree, const ut_prp_slot* slots,
	uint32_t nslots, const uint32_t* kring, ut_prp_summary* out, uint32_t* raw_hits)
{
	if (!p || !face_tree || !edge_tree || (!slots && nslots) || !kring)
		return ut::kInvalid;
	return guarded(p->r.ctx,
		[&]
		{
			std::vector<ut::PrpSlot> sl(nslots);
			std::memcpy(sl.data(), slots, sizeof(ut::PrpSlot) * nslots);
			for (const auto& q : sl)
				UT_REQUIRE(q.contour < p->r.batch_n, "prp: slot contour not in the prepared batch");
			std::vector<uint32_t>	   kr(kring, kring + 2 * p->r.batch_n);
			std::vector<ut::PrpSummary> res;
			const uint32_t			   nh = p->r.eval_slots(sl, face_tree->b, edge_tree->b, kr, res);
			if (out)
				std::memcpy(out, res.data(), sizeof(ut::PrpSummary) * res.size());
			if (raw_hits)
				*raw_hits = nh;
		});
}

UT_API int ut_prp_hits(ut_prp* p, ut_prp_hit* out, uint32_t cap)
{
	if (!p || (!out && cap))
		return ut::kInvalid;
	return guarded(p->r.ctx,
		[&]
		{
			const uint32_t n = std::min(cap, p->r.last_raw_hits);
			if (n)
			{
				UT_HIP_CHECK(hipMemcpyAsync(out, p->r.hits_sorted.get(), sizeof(ut::PrpHit) * n, hipMemcpyDeviceToHost,
					p->r.ctx->stream));
				UT_HIP_CHECK(hipStreamSynchronize(p->r.ctx->stream));
			}
		});
}

UT_API int ut_prp_normals(ut_prp* p, float* out)
{
	if (!p || !out)
		return ut::kInvalid;
	return guarded(p->r.ctx,
		[&]
		{
			const size_t n = 3 * ((size_t)p->r.nf + p->r.ne + p->r.nv);
			UT_HIP_CHECK(hipMemcpyAsync(out, p->r.normals.get(), 4 * n, hipMemcpyDeviceToHost, p->r.ctx->stream));
			UT_HIP_CHECK(hipStreamSynchronize(p->r.ctx->stream));
		});
}

UT_API int ut_solver_profile(ut_solver* s, int what, uint32_t reps, float* avg_ms)
{
	if (!s || !avg_ms || what != UT_PROFILE_PCG_SPMV)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			const bool fp64 = s->s.p.pcg.precision == ut::kPcgFp64;
			*avg_ms = s->s.pcg.profile_spmv(fp64 ? (const void*)s->s.vals64.get() : (const void*)s->s.vals32.get(),
				!fp64, reps);
		});
}

UT_API int ut_solver_get_positions(ut_solver* s, int which, float* x)
{
	if (!s || !x || which < 0 || which > 2)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			const float* src = which == 0 ? s->s.x.get() : which == 1 ? s->s.x_iter_start.get() : s->s.x_step_start.get();
			UT_HIP_CHECK(
				hipMemcpyAsync(x, src, 12 * (size_t)s->s.d.nv, hipMemcpyDeviceToDevice, s->s.ctx->stream));
		});
}

UT_API int ut_solver_get_q(ut_solver* s, double* q, double* vq)
{
	if (!s || (!q && !vq))
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			UT_REQUIRE(s->s.abd_mode(), "solver: no affine bodies");
			const size_t bytes = 96 * (size_t)s->s.d.abd.nb;
			if (q)
				UT_HIP_CHECK(hipMemcpyAsync(q, s->s.q.get(), bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
			if (vq)
				UT_HIP_CHECK(hipMemcpyAsync(vq, s->s.vq.get(), bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
		});
}

UT_API int ut_solver_set_q(ut_solver* s, const double* q, const double* vq)
{
	if (!s || (!q && !vq))
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			UT_REQUIRE(s->s.abd_mode(), "solver: no affine bodies");
			const size_t bytes = 96 * (size_t)s->s.d.abd.nb;
			if (q)
			{
				UT_HIP_CHECK(hipMemcpyAsync(s->s.q.get(), q, bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
				ut::abd_q_to_x(s->s.ctx, s->s.d.abd, s->s.q.get(), s->s.x.get(), nullptr);
			}
			if (vq)
				UT_HIP_CHECK(hipMemcpyAsync(s->s.vq.get(), vq, bytes, hipMemcpyDeviceToDevice, s->s.ctx->stream));
		});
}

UT_API int ut_solver_responses(ut_solver* s, void* records, uint32_t* meta, uint32_t cap, uint32_t* count)
{
	if (!s || !count)
		return ut::kInvalid;
	return guarded(s->s.ctx,
		[&]
		{
			const uint32_t n = s->s.last_responses;
			*count = n;
			const uint32_t k = std::min(n, cap);
			if (k == 0)
				return;
			UT_REQUIRE(records && meta, "null destination");
			const hipStream_t st = s->s.ctx->stream;
			UT_HIP_CHECK(hipMemcpyAsync(records, s->s.prp.responses.get(), sizeof(ut::ContactRec) * k,
				hipMemcpyDeviceToDevice, st));
			UT_HIP_CHECK(hipMemcpyAsync(meta, s->s.prp.response_meta.get(), 16 * (size_t)k, hipMemcpyDeviceToDevice, st));
		});
}

UT_API int ut_solver_pcg_resolve(ut_solver* s, int precision, ut_pcg_report* report, float* ms)
{
	if (!s || !report)
		return ut::kInvalid;
	int conv = 1;
	const int rc = guarded(s->s.ctx,
		[&]
		{
			const ut::PcgReport r = s->s.resolve_last_system(precision, ms);
			report->relative_residual = r.relative_residual;
			report->infinity_norm = r.infinity_norm;
			report->iterations = r.iterations;
			report->refinements = r.refinements;
			report->converged = r.converged;
			report->failure = r.failure;
			conv = r.converged;
		});
	return rc != ut::kOk ? rc : (conv ? ut::kOk : ut::kNotConverged);
}
