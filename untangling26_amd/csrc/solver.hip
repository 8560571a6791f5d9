// NewtonSolver::physics_step_GPU on gfx950 (newton_solver.cpp:2482-2767):
// LBVH rebuild, prediction, then per Newton iteration: refit + fused VF/EE
// proximity DCD + fused EF intersection detection + contours, assembly of the
// incremental potential, PCG (FP64 or FP32 with FP64 refinement), and the
// global additive-CCD line search; finally the velocity update.
//
// Host synchronisations per iteration: one count readback after detection
// (resize-and-retry when a list overflowed, newton_solver.cpp:1402-1438), the
// contour build's two, and the PCG convergence polls.  The line-search step
// x = x0 + 0.9 toi dx reads toi on the device.
#include "ut_common.hpp"
#include "solver.hpp"

#include <cmath>

namespace ut
{

__global__ void k_predict(uint32_t nv, const float* __restrict__ x, const float* __restrict__ v, double h, double gx,
	double gy, double gz, const uint32_t* __restrict__ vprop, int quasi_static, double* __restrict__ x_tilde)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nv)
		return;
	// fn_predict_position (newton_solver.cpp:561-598): v_pred = h g (+ v unless quasi-static), x_tilde = x + h v_pred
	const bool fixed = vprop && (vprop[i] & kPropFixed);
	const double g[3] = { gx, gy, gz };
	for (int a = 0; a < 3; ++a)
	{
		const double vp = (quasi_static ? 0.0 : (double)v[3 * i + a]) + (fixed ? 0.0 : h * g[a]);
		x_tilde[3 * i + a] = (double)x[3 * i + a] + (fixed ? 0.0 : h * vp);
	}
}

__global__ void k_negate(size_t n, const double* __restrict__ a, double* __restrict__ b)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		b[i] = -a[i];
}

__global__ void k_trial(size_t n, const float* __restrict__ x0, const double* __restrict__ dx, float* __restrict__ xe)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		xe[i] = (float)((double)x0[i] + dx[i]);
}

// x = x0 + alpha dx with alpha = toi, scaled by 0.9 when an impact limited it
// (fn_scaling_toi, newton_solver.cpp:191-202, 1481-1482).
__global__ void k_apply_step(size_t n, const float* __restrict__ x0, const double* __restrict__ dx,
	const float* __restrict__ toi, float* __restrict__ x)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float t = toi ? *toi : 1.0f;
	if (t != 1.0f)
		t = 0.9f * t;
	x[i] = (float)((double)x0[i] + (double)t * dx[i]);
}

__global__ void k_update_velocity(size_t n, const float* __restrict__ x, const float* __restrict__ x_start, double h,
	double damping, float* __restrict__ v)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		v[i] = (float)(((double)x[i] - (double)x_start[i]) / h * exp(-damping * h)); // fn_update_velocity
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0u;
}

// affine bodies: q_tilde = q + h v_q + h^2 g on the translation (fn_predict_position's
// affine branch, newton_solver.cpp:600-640; quasi-static drops v_q)
__global__ void k_abd_predict(uint32_t nb, const double* __restrict__ q, const double* __restrict__ vq, double h,
	double gx, double gy, double gz, int quasi_static, double* __restrict__ q_tilde)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= 12 * nb)
		return;
	const uint32_t k = i % 12;
	const double   g = k == 0 ? gx : k == 1 ? gy : k == 2 ? gz : 0.0;
	const double   vp = (quasi_static ? 0.0 : vq[i]) + h * g;
	q_tilde[i] = q[i] + h * vp;
}

__global__ void k_abd_trial(uint32_t n, const double* __restrict__ q0, const double* __restrict__ dq,
	const float* __restrict__ toi, double* __restrict__ q)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double t = 1.0;
	if (toi)
	{
		const float f = *toi;
		t = f != 1.0f ? 0.9 * (double)f : 1.0;
	}
	q[i] = q0[i] + t * dq[i];
}

__global__ void k_abd_velocity(uint32_t n, const double* __restrict__ q, const double* __restrict__ q0, double h,
	double damping, double* __restrict__ vq)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		vq[i] = (q[i] - q0[i]) / h * exp(-damping * h);
}

__global__ void k_neg(uint32_t n, const double* __restrict__ a, double* __restrict__ b)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		b[i] = -a[i];
}

static uint32_t g(size_t n) { return (uint32_t)((n + 255) / 256); }

void NewtonSolver::init(Context* c, const SolverDesc& desc, const SolverParams& params, const float* x0, const float* v0)
{
	ctx = c;
	d = desc;
	p = params;
	UT_REQUIRE(d.nv && d.nf && d.ne && d.edges && d.faces && d.mass && d.d_hat && d.offset, "solver: missing mesh data");
	if (abd_mode())
	{
		UT_REQUIRE(d.abd.nv == d.nv && d.abd.body && d.abd.xbar && d.abd.mass4 && d.abd.volume && d.abd.kappa && d.q0,
			"solver: affine bodies need every vertex's body, model position, the mass matrices and q0");
		UT_REQUIRE(d.vprop, "solver: affine bodies need the rigid property word of every vertex (kPropRigid | body << 8)");
	}
	else
		UT_REQUIRE(d.slice_off && d.col && d.diag_slot && d.face_slots, "solver: missing block pattern");
	const size_t n3 = 3 * (size_t)d.nv;
	x.reserve(n3);
	v.reserve(n3);
	x_step_start.reserve(n3);
	x_iter_start.reserve(n3);
	x_end.reserve(n3);
	x_tilde.reserve(n3);
	grad.reserve(n3);
	rhs.reserve(n3);
	dx.reserve(n3);
	vals64.reserve(9 * (size_t)d.nslots);
	vals32.reserve(9 * (size_t)d.nslots);
	energy.reserve(4);
	toi.reserve(1);
	counts.reserve(8);
	UT_HIP_CHECK(hipMemcpyAsync(x.get(), x0, 4 * n3, hipMemcpyDeviceToDevice, ctx->stream));
	if (v0)
		UT_HIP_CHECK(hipMemcpyAsync(v.get(), v0, 4 * n3, hipMemcpyDeviceToDevice, ctx->stream));
	else
		UT_HIP_CHECK(hipMemsetAsync(v.get(), 0, 4 * n3, ctx->stream));
	face_tree.init(ctx, kTreeFace, d.nf);
	edge_tree.init(ctx, kTreeEdge, d.ne);
	if (abd_mode())
	{
		const size_t nq = 12 * (size_t)d.abd.nb;
		for (DevBuf<double>* b : { &q, &vq, &q_step_start, &q_iter_start, &q_tilde, &gq, &rhs_q, &dq, &q_trial })
			b->reserve(nq);
		blocks.reserve(144 * (size_t)d.abd.nb);
		abd_energy.reserve(2);
		UT_HIP_CHECK(hipMemcpyAsync(q.get(), d.q0, 8 * nq, hipMemcpyDeviceToDevice, ctx->stream));
		UT_HIP_CHECK(hipMemsetAsync(vq.get(), 0, 8 * nq, ctx->stream));
		abd_q_to_x(ctx, d.abd, q.get(), x.get(), nullptr);
		abd_pcg.init(ctx, d.abd.nb);
	}
	else
	{
		pcg.init(ctx, d.nv, d.nslices, d.nslots, d.slice_off, d.col, d.diag_slot);
		build_colors();
	}
	if (d.face_adj_edges && d.edge_adj_faces)
	{
		contours.init(ctx, d.nv, d.ne, d.nf, d.edges, d.faces, d.face_adj_edges, d.edge_adj_faces);
		if (!p.ignore_near_zero_dist_pairs)
			build_vert_adj_edges();
		if (d.rest_vert_area && d.rest_edge_area && d.rest_face_area)
		{
			if (!d.rest_x)
			{
				rest_copy.reserve(n3);
				UT_HIP_CHECK(hipMemcpyAsync(rest_copy.get(), x0, 4 * n3, hipMemcpyDeviceToDevice, ctx->stream));
				d.rest_x = rest_copy.get();
			}
			prp.init(ctx, d.nv, d.ne, d.nf, d.edges, d.faces, d.face_adj_edges, d.edge_adj_faces, d.rest_x,
				d.rest_vert_area, d.rest_edge_area, d.rest_face_area, d.mass, p.prp);
			regions.init(ctx, d.nv);
		}
	}
	ensure_contact_capacity(std::max<uint32_t>(4096, 4 * d.nv), std::max<uint32_t>(1024, d.ne / 4));
	UT_HIP_CHECK(hipHostMalloc((void**)&h_counts, 32, hipHostMallocDefault));
	for (auto& e : ev)
		UT_HIP_CHECK(hipEventCreate(&e));
}

void NewtonSolver::build_colors()
{
	const hipStream_t	  s = ctx->stream;
	std::vector<uint32_t> f(3 * (size_t)d.nf), h(4 * (size_t)d.nh), fo, ho;
	UT_HIP_CHECK(hipMemcpyAsync(f.data(), d.faces, 12 * (size_t)d.nf, hipMemcpyDeviceToHost, s));
	if (d.nh && d.hinges)
		UT_HIP_CHECK(hipMemcpyAsync(h.data(), d.hinges, 16 * (size_t)d.nh, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	// more than 256 classes (a vertex in hundreds of elements): atomic assembly
	colored = color_elements(f.data(), d.nf, 3, d.nv, fo, face_color_off)
		&& color_elements(h.data(), d.hinges ? d.nh : 0u, 4, d.nv, ho, hinge_color_off);
	if (!colored)
		return;
	face_order.reserve(std::max<size_t>(fo.size(), 1));
	hinge_order.reserve(std::max<size_t>(ho.size(), 1));
	UT_HIP_CHECK(hipMemcpyAsync(face_order.get(), fo.data(), 4 * fo.size(), hipMemcpyHostToDevice, s));
	if (!ho.empty())
		UT_HIP_CHECK(hipMemcpyAsync(hinge_order.get(), ho.data(), 4 * ho.size(), hipMemcpyHostToDevice, s));
	UT_HIP_CHECK(hipStreamSynchronize(s));
}

void NewtonSolver::build_vert_adj_edges()
{
	std::vector<uint32_t> e(2 * (size_t)d.ne), off(d.nv + 1, 0), col(2 * (size_t)d.ne);
	if (d.ne)
	{
		UT_HIP_CHECK(hipMemcpyAsync(e.data(), d.edges, 8 * (size_t)d.ne, hipMemcpyDeviceToHost, ctx->stream));
		UT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	}
	for (uint32_t v : e)
	{
		UT_REQUIRE(v < d.nv, "solver: edge vertex out of range");
		off[v + 1]++;
	}
	for (uint32_t v = 0; v < d.nv; ++v)
		off[v + 1] += off[v];
	std::vector<uint32_t> fill(off.begin(), off.end() - 1);
	for (uint32_t k = 0; k < d.ne; ++k) // ascending edge ids per vertex
		for (int j = 0; j < 2; ++j)
			col[fill[e[2 * k + j]]++] = k;
	vert_adj_off.reserve(d.nv + 1);
	vert_adj_edges.reserve(col.size());
	UT_HIP_CHECK(hipMemcpyAsync(vert_adj_off.get(), off.data(), 4 * off.size(), hipMemcpyHostToDevice, ctx->stream));
	if (!col.empty())
		UT_HIP_CHECK(
			hipMemcpyAsync(vert_adj_edges.get(), col.data(), 4 * col.size(), hipMemcpyHostToDevice, ctx->stream));
	UT_HIP_CHECK(hipStreamSynchronize(ctx->stream)); // the host vectors go out of scope
}

NewtonSolver::~NewtonSolver()
{
	if (h_counts)
		(void)hipHostFree(h_counts);
	for (auto& e : ev)
		if (e)
			(void)hipEventDestroy(e);
}

void NewtonSolver::ensure_contact_capacity(uint32_t need_contacts, uint32_t need_ef, uint32_t keep)
{
	if (need_contacts > cap_contacts)
	{
		// grow, preserving the first `keep` records
		DevBuf<ContactRec> np;
		DevBuf<uint32_t>   ni;
		np.reserve(need_contacts);
		ni.reserve(2 * (size_t)need_contacts);
		if (keep)
		{
			UT_HIP_CHECK(hipMemcpyAsync(np.get(), contact_pairs.get(), sizeof(ContactRec) * keep,
				hipMemcpyDeviceToDevice, ctx->stream));
			UT_HIP_CHECK(hipMemcpyAsync(ni.get(), contact_ids.get(), 8 * (size_t)keep, hipMemcpyDeviceToDevice,
				ctx->stream));
			UT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		}
		std::swap(contact_pairs.ptr, np.ptr), std::swap(contact_pairs.cap, np.cap);
		std::swap(contact_ids.ptr, ni.ptr), std::swap(contact_ids.cap, ni.cap);
		cap_contacts = need_contacts;
	}
	if (need_ef > cap_ef)
	{
		cap_ef = need_ef;
		ef_pairs.reserve(cap_ef);
		ef_ids.reserve(2 * (size_t)cap_ef);
	}
}

// device_update_contact_list (newton_solver.cpp:1393-1445) with fused queries
void NewtonSolver::detect(StepReport& rep)
{
	hipStream_t s = ctx->stream;
	face_tree.update_leaves_and_refit(d.offset, x.get(), x.get(), d.faces);
	edge_tree.update_leaves_and_refit(d.offset, x.get(), x.get(), d.edges);
	for (int attempt = 0; attempt < 10; ++attempt)
	{
		k_zero_u32<<<1, 64, 0, s>>>(counts.get(), 8);
		ContactArgs a;
		a.num_verts = d.nv;
		a.num_edges = d.ne;
		a.x = x.get();
		a.edges = d.edges;
		a.faces = d.faces;
		a.vprop = d.vprop;
		a.d_hat = d.d_hat;
		a.offset = d.offset;
		a.rest_vert_area = d.rest_vert_area;
		a.rest_edge_area = d.rest_edge_area;
		a.rest_face_area = d.rest_face_area;
		a.kappa = p.kappa;
		a.energy = p.energy;
		a.count = counts.get() + 3; // VF and EE share one list
		a.pairs = contact_pairs.get();
		a.pair_ids = reinterpret_cast<uint2*>(contact_ids.get());
		a.max_pairs = cap_contacts;
		contact_query(ctx, 0, a, &face_tree, nullptr, nullptr, 0);
		UT_HIP_CHECK(hipMemcpyAsync(counts.get() + 0, counts.get() + 3, 4, hipMemcpyDeviceToDevice, s));
		contact_query(ctx, 1, a, &edge_tree, nullptr, nullptr, 0);
		if (p.use_untangling)
		{
			EfDcdArgs e;
			e.x = x.get();
			e.edges = d.edges;
			e.faces = d.faces;
			e.vprop = d.vprop;
			e.d_hat = d.d_hat;
			e.count = counts.get() + 2;
			e.pairs = reinterpret_cast<uint32_t*>(ef_pairs.get());
			e.pair_ids = ef_ids.get();
			e.max_pairs = cap_ef;
			ef_detect(ctx, face_tree, e, d.ne, d.offset);
			if (!p.ignore_near_zero_dist_pairs && contours.ctx)
			{
				// fn_add_ef_from_proximity_pairs over the VF + EE proximity list
				EfFromProximityArgs z;
				z.recs = reinterpret_cast<const uint32_t*>(contact_pairs.get());
				z.rec_ids = contact_ids.get();
				z.rec_count = counts.get() + 3;
				z.rec_cap = cap_contacts;
				z.x = x.get();
				z.edges = d.edges;
				z.faces = d.faces;
				z.vert_adj_off = vert_adj_off.get();
				z.vert_adj_edges = vert_adj_edges.get();
				z.edge_adj_faces = d.edge_adj_faces;
				z.count = counts.get() + 2;
				z.pairs = reinterpret_cast<uint32_t*>(ef_pairs.get());
				z.pair_ids = ef_ids.get();
				z.max_pairs = cap_ef;
				ef_from_proximity(ctx, z);
			}
		}
		UT_HIP_CHECK(hipMemcpyAsync(h_counts, counts.get(), 16, hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		const uint32_t nc = h_counts[3], nef = h_counts[2];
		if (nc <= cap_contacts && nef <= cap_ef)
		{
			rep.vf_pairs = h_counts[0];
			rep.ee_pairs = nc - h_counts[0];
			rep.ef_pairs = nef;
			// canonical record order: (kind, query, leaf)
			const uint32_t mq = std::max(d.nv, d.ne), ml = std::max(d.nf, d.ne);
			const int	   qb = std::max(1, 32 - __builtin_clz(std::max(mq, 2u) - 1));
			const int	   lb = std::max(1, 32 - __builtin_clz(std::max(ml, 2u) - 1));
			csort.sort(ctx, contact_pairs.get(), contact_ids.get(), nc, h_counts[0], qb, lb);
			return;
		}
		ensure_contact_capacity(nc > cap_contacts ? 2 * nc : cap_contacts, nef > cap_ef ? 2 * nef : cap_ef);
	}
	throw Error(kCapacity, "contact lists still out of range after 10 resize attempts");
}

// device_ccd_line_search (newton_solver.cpp:1446-1483): the global time of impact
// from x_iter_start to x_end over the swept trees, pairs across a contour skipped
void NewtonSolver::ccd_line_search()
{
	face_tree.update_leaves_and_refit(d.offset, x_iter_start.get(), x_end.get(), d.faces);
	edge_tree.update_leaves_and_refit(d.offset, x_iter_start.get(), x_end.get(), d.edges);
	ContactArgs c;
	c.num_verts = d.nv;
	c.num_edges = d.ne;
	c.x = x_iter_start.get();
	c.x_end = x_end.get();
	c.edges = d.edges;
	c.faces = d.faces;
	c.vprop = d.vprop;
	c.d_hat = d.d_hat;
	c.offset = d.offset;
	c.toi = toi.get();
	c.filter = regions.view();
	contact_query(ctx, 2, c, &face_tree, nullptr, nullptr, 0);
	contact_query(ctx, 3, c, &edge_tree, nullptr, nullptr, 0);
}

StepReport NewtonSolver::physics_step_GPU()
{
	StepReport	   rep;
	hipStream_t	   s = ctx->stream;
	const size_t   n3 = 3 * (size_t)d.nv;
	const double   h = p.dt;
	UT_HIP_CHECK(hipEventRecord(ev[0], s));
	UT_HIP_CHECK(hipMemcpyAsync(x_step_start.get(), x.get(), 4 * n3, hipMemcpyDeviceToDevice, s));
	if (p.use_self_collision)
	{
		face_tree.reduce_aabb(x.get(), d.faces);
		face_tree.construct_tree();
		edge_tree.reduce_aabb(x.get(), d.edges);
		edge_tree.construct_tree();
	}
	if (abd_mode())
	{
		const uint32_t nq = 12 * d.abd.nb;
		UT_HIP_CHECK(hipMemcpyAsync(q_step_start.get(), q.get(), 8 * (size_t)nq, hipMemcpyDeviceToDevice, s));
		k_abd_predict<<<g(nq), 256, 0, s>>>(d.abd.nb, q.get(), vq.get(), h, p.gravity[0], p.gravity[1],
			p.gravity[2], p.quasi_static, q_tilde.get());
	}
	else
		k_predict<<<g(d.nv), 256, 0, s>>>(d.nv, x.get(), v.get(), h, p.gravity[0], p.gravity[1], p.gravity[2],
			d.vprop, p.quasi_static, x_tilde.get());
	const bool fp64 = p.pcg.precision == kPcgFp64;
	float	   t_det = 0, t_ct = 0, t_asm = 0, t_pcg = 0, t_ccd = 0, t_prp = 0;
	for (int it = 0; it < std::max(1, p.nonlinear_iters); ++it)
	{
		UT_HIP_CHECK(hipMemcpyAsync(x_iter_start.get(), x.get(), 4 * n3, hipMemcpyDeviceToDevice, s));
		if (abd_mode())
			UT_HIP_CHECK(hipMemcpyAsync(q_iter_start.get(), q.get(), 96 * (size_t)d.abd.nb, hipMemcpyDeviceToDevice, s));
		UT_HIP_CHECK(hipEventRecord(ev[1], s));
		uint32_t ncontacts = 0;
		if (p.use_self_collision)
		{
			detect(rep);
			ncontacts = rep.vf_pairs + rep.ee_pairs;
		}
		UT_HIP_CHECK(hipEventRecord(ev[2], s));
		regions.clear();
		last_responses = 0;
		if (p.use_self_collision && p.use_untangling && contours.ctx)
		{
			contours.build(rep.ef_pairs, ef_pairs.get(), ef_ids.get(), !p.ignore_near_zero_dist_pairs);
			rep.ef_total = contours.num_total;
			rep.contours = contours.num_contours;
			if (contours.num_total > 0 && prp.ctx)
			{
				// host_resolve -> PRP responses appended after the proximity pairs
				// (upload_response_pairs_to_gpu, intersection_resolver2.cpp:2223-2294)
				uint32_t nresp = 0;
				if (p.use_prp && contours.num_contours > 0)
				{
					PrpStats ps;
					nresp = prp.resolve(x.get(), contours, face_tree, edge_tree, last_choices, &ps);
					rep.prp_responses += nresp;
					last_responses = nresp;
					rep.prp_contours += ps.contours;
					rep.prp_levels += ps.levels;
					rep.prp_hits += ps.hits;
					rep.prp_slots += ps.slots;
					for (int k = 0; k < 3; ++k)
						rep.prp_rays[k] += ps.rays[k];
					rep.prp_ee.launches += ps.ee.launches, rep.prp_ee.ms += ps.ee.ms;
					rep.prp_ee.requests += ps.ee.requests, rep.prp_ee.visits += ps.ee.visits;
					rep.prp_ee.leaf_tests += ps.ee.leaf_tests, rep.prp_ee.hits += ps.ee.hits;
					rep.prp_ee.cand_tests += ps.ee.cand_tests, rep.prp_ee.wave_steps += ps.ee.wave_steps;
					t_prp += ps.ms_prepare + ps.ms_levels + ps.ms_final;
				}
				else
					prp.last_final_hits = 0;
				// filter_invalid_pairs (intersection_resolver.cpp:2640-2682) on the proximity pairs
				if (p.region_filter)
				{
					regions.build(prp, contours);
					regions.filter_proximity(contact_pairs.get(), ncontacts, x.get(), d.offset);
				}
				if (nresp)
				{
					ensure_contact_capacity(ncontacts + nresp, cap_ef, ncontacts);
					UT_HIP_CHECK(hipMemcpyAsync(contact_pairs.get() + ncontacts, prp.responses.get(),
						sizeof(ContactRec) * nresp, hipMemcpyDeviceToDevice, s));
					h_counts[4] = ncontacts + nresp;
					UT_HIP_CHECK(hipMemcpyAsync(counts.get() + 3, h_counts + 4, 4, hipMemcpyHostToDevice, s));
					UT_HIP_CHECK(hipStreamSynchronize(s));
					ncontacts += nresp;
				}
			}
		}
		// fn_process_collision_pair_friction (narrow_phase.cpp:1455-1521, run before every
		// assembly): the first iteration's proximity records get mu_lambda = -mu k1 (after the
		// region filter, so disabled records get none) and are kept for the later iterations,
		// which append them with their relative motion since the step start and no repulsion
		// (the reference keeps them at the front of the list, reset_narrowphase_count).
		const uint32_t nprox = rep.vf_pairs + rep.ee_pairs;
		if (d.friction_mu && it == 0)
		{
			nfric = 0;
			if (p.use_self_collision && nprox > 0)
			{
				h_counts[5] = nprox;
				UT_HIP_CHECK(hipMemcpyAsync(counts.get() + 5, h_counts + 5, 4, hipMemcpyHostToDevice, s));
				friction_init(ctx, contact_pairs.get(), counts.get() + 5, nprox, d.friction_mu);
				if (p.nonlinear_iters > 1)
				{
					fric_pairs.reserve(nprox);
					UT_HIP_CHECK(hipMemcpyAsync(fric_pairs.get(), contact_pairs.get(), sizeof(ContactRec) * nprox,
						hipMemcpyDeviceToDevice, s));
					nfric = nprox;
				}
			}
		}
		else if (d.friction_mu && nfric > 0)
		{
			ensure_contact_capacity(ncontacts + nfric, cap_ef, ncontacts);
			// detect() cleared the counters of this iteration: the lagged kernel reads its
			// record count from counts[5], so it is set again here (all nfric records written)
			h_counts[5] = nfric;
			UT_HIP_CHECK(hipMemcpyAsync(counts.get() + 5, h_counts + 5, 4, hipMemcpyHostToDevice, s));
			friction_lagged(ctx, fric_pairs.get(), counts.get() + 5, nfric, x.get(), nullptr, x_step_start.get(),
				nullptr, contact_pairs.get() + ncontacts);
			ncontacts += nfric;
			h_counts[4] = ncontacts;
			UT_HIP_CHECK(hipMemcpyAsync(counts.get() + 3, h_counts + 4, 4, hipMemcpyHostToDevice, s));
			UT_HIP_CHECK(hipStreamSynchronize(s));
		}
		UT_HIP_CHECK(hipEventRecord(ev[3], s));
		if (abd_mode())
		{
			// q-space assembly and solve (the affine-body branch of the Newton step)
			const uint32_t	nq = 12 * d.abd.nb;
			AbdAssembleArgs aa;
			aa.q = q.get();
			aa.q_tilde = q_tilde.get();
			aa.inv_h2 = 1.0 / (h * h);
			aa.inertia_grad_scale = p.quasi_static ? 0.0 : 1.0;
			aa.inertia_hess_scale = (p.quasi_static && !p.qs_inertia_hessian) ? 0.0 : 1.0;
			aa.contacts = ncontacts ? contact_pairs.get() : nullptr;
			aa.contact_count = counts.get() + 3;
			aa.contact_cap = ncontacts ? cap_contacts : 0u;
			aa.grad = gq.get();
			aa.blocks = blocks.get();
			aa.energy = abd_energy.get();
			abd_assemble(ctx, d.abd, aa);
			k_neg<<<g(nq), 256, 0, s>>>(nq, gq.get(), rhs_q.get());
			UT_HIP_CHECK(hipEventRecord(ev[4], s));
			const AbdPcgReport pr = abd_pcg.solve(d.abd, blocks.get(), ncontacts ? contact_pairs.get() : nullptr,
				ncontacts ? counts.get() + 3 : nullptr, ncontacts ? cap_contacts : 0u, rhs_q.get(), dq.get(),
				p.pcg.precision == kPcgFp64 ? kPcgFp64 : kPcgMixed, p.pcg.max_iters, p.pcg.rel_tol);
			rep.pcg_iterations += pr.iterations;
			rep.pcg_refinements += pr.refinements;
			rep.pcg_relative_residual = pr.relative_residual;
			rep.pcg_converged = pr.converged;
			UT_HIP_CHECK(hipEventRecord(ev[5], s));
			// trial positions x(q + dq), CCD, then q += alpha dq and x = x(q)
			k_abd_trial<<<g(nq), 256, 0, s>>>(nq, q_iter_start.get(), dq.get(), nullptr, q_trial.get());
			abd_q_to_x(ctx, d.abd, q_trial.get(), x_end.get(), nullptr);
			toi_reset(ctx, toi.get());
			if (p.use_self_collision && p.use_ccd)
				ccd_line_search();
			k_abd_trial<<<g(nq), 256, 0, s>>>(nq, q_iter_start.get(), dq.get(), toi.get(), q.get());
			abd_q_to_x(ctx, d.abd, q.get(), x.get(), nullptr);
			UT_HIP_CHECK(hipEventRecord(ev[6], s));
			UT_HIP_CHECK(hipEventSynchronize(ev[6]));
			float ms;
			UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[1], ev[2]));
			t_det += ms;
			UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[2], ev[3]));
			t_ct += ms;
			UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[3], ev[4]));
			t_asm += ms;
			UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[4], ev[5]));
			t_pcg += ms;
			UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[5], ev[6]));
			t_ccd += ms;
			has_system = true;
			continue;
		}
		AssembleArgs a;
		a.nv = d.nv;
		a.nf = d.nf;
		a.nh = d.nh;
		a.nslots = d.nslots;
		a.x = x.get();
		a.x_tilde = x_tilde.get();
		a.mass = d.mass;
		a.inv_h2 = 1.0 / (h * h);
		// quasi-static: no inertia gradient/energy; HessianOnly keeps the diagonal (soft_inertia_energy.cpp:68-80)
		a.inertia_grad_scale = p.quasi_static ? 0.0 : 1.0;
		a.inertia_hess_scale = (p.quasi_static && !p.qs_inertia_hessian) ? 0.0 : 1.0;
		a.faces = d.faces;
		a.dm = d.dm;
		a.face_area = d.rest_face_area;
		a.mu_lambda = d.mu_lambda;
		a.face_slots = d.face_slots;
		a.hinges = d.hinges;
		a.rest_angle = d.rest_angle;
		a.bend_stiff_area = d.bend_stiff_area;
		a.hinge_slots = d.hinge_slots;
		a.contacts = ncontacts ? contact_pairs.get() : nullptr;
		a.contact_count = counts.get() + 3;
		a.contact_cap = cap_contacts;
		a.diag_slot = d.diag_slot;
		a.grad = grad.get();
		a.vals = vals64.get();
		a.vals32 = fp64 ? nullptr : vals32.get();
		a.energy = energy.get();
		if (colored)
		{
			a.face_order = face_order.get();
			a.face_color_off = face_color_off.data();
			a.face_colors = (uint32_t)face_color_off.size() - 1;
			a.hinge_order = hinge_order.get();
			a.hinge_color_off = hinge_color_off.data();
			a.hinge_colors = (uint32_t)hinge_color_off.size() - 1;
		}
		// the PCG's per-vertex contact lists, built once and shared with the assembly
		pcg.contacts = ncontacts ? contact_pairs.get() : nullptr;
		pcg.contact_count = counts.get() + 3;
		pcg.contact_cap = cap_contacts;
		pcg.prepare_contacts();
		pcg.contacts_prepared = true;
		if (pcg.ct_entries && pcg.ct_sorted)
		{
			a.ct_off = pcg.ct_off.get();
			a.ct_entries = pcg.ct_sorted;
		}
		assemble(ctx, a);
		k_negate<<<g(n3), 256, 0, s>>>(n3, grad.get(), rhs.get());
		UT_HIP_CHECK(hipEventRecord(ev[4], s));
		const PcgReport pr =
			pcg.solve(p.pcg, fp64 ? (const void*)vals64.get() : (const void*)vals32.get(), rhs.get(), dx.get(),
				p.pcg.precision == kPcgMixed ? vals64.get() : nullptr);
		rep.pcg_iterations += pr.iterations;
		rep.pcg_refinements += pr.refinements;
		rep.pcg_relative_residual = pr.relative_residual;
		rep.pcg_converged = pr.converged;
		has_system = true;
		UT_HIP_CHECK(hipEventRecord(ev[5], s));
		// device_ccd_line_search (newton_solver.cpp:1446-1483)
		k_trial<<<g(n3), 256, 0, s>>>(n3, x_iter_start.get(), dx.get(), x_end.get());
		toi_reset(ctx, toi.get());
		if (p.use_self_collision && p.use_ccd)
			ccd_line_search();
		k_apply_step<<<g(n3), 256, 0, s>>>(n3, x_iter_start.get(), dx.get(), toi.get(), x.get());
		UT_HIP_CHECK(hipEventRecord(ev[6], s));
		UT_HIP_CHECK(hipEventSynchronize(ev[6]));
		float ms;
		UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[1], ev[2]));
		t_det += ms;
		UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[2], ev[3]));
		t_ct += ms;
		UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[3], ev[4]));
		t_asm += ms;
		UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[4], ev[5]));
		t_pcg += ms;
		UT_HIP_CHECK(hipEventElapsedTime(&ms, ev[5], ev[6]));
		t_ccd += ms;
	}
	k_update_velocity<<<g(n3), 256, 0, s>>>(n3, x.get(), x_step_start.get(), h, p.damping, v.get());
	if (abd_mode())
		k_abd_velocity<<<g(12 * (size_t)d.abd.nb), 256, 0, s>>>(12 * d.abd.nb, q.get(), q_step_start.get(), h,
			p.damping, vq.get());
	UT_HIP_CHECK(hipEventRecord(ev[7], s));
	UT_HIP_CHECK(hipMemcpyAsync(&rep.toi, toi.get(), 4, hipMemcpyDeviceToHost, s));
	if (abd_mode())
	{
		rep.energy[2] = 0.0;
		UT_HIP_CHECK(hipMemcpyAsync(rep.energy, abd_energy.get(), 16, hipMemcpyDeviceToHost, s));
	}
	else
		UT_HIP_CHECK(hipMemcpyAsync(rep.energy, energy.get(), 24, hipMemcpyDeviceToHost, s));
	UT_HIP_CHECK(hipEventSynchronize(ev[7]));
	UT_HIP_CHECK(hipStreamSynchronize(s));
	float total;
	UT_HIP_CHECK(hipEventElapsedTime(&total, ev[0], ev[7]));
	rep.ms_detect = t_det;
	rep.ms_contours = t_ct;
	rep.ms_assemble = t_asm;
	rep.ms_pcg = t_pcg;
	rep.ms_ccd = t_ccd;
	rep.ms_total = total;
	rep.ms_prp = t_prp;
	return rep;
}

void NewtonSolver::set_params(const SolverParams& np)
{
	UT_REQUIRE(np.prp.total_dirs >= 1 && np.prp.total_dirs <= 64, "solver: PRP directions must be in [1, 64]");
	UT_REQUIRE(np.prp.batch_contours >= 1, "solver: PRP batch_contours must be positive");
	UT_REQUIRE(np.pcg.check_every >= 1, "solver: PCG check_every must be >= 1");
	if (!np.ignore_near_zero_dist_pairs && contours.ctx && !vert_adj_off.get())
		build_vert_adj_edges();
	p = np;
	prp.params = p.prp;
}

PcgReport NewtonSolver::resolve_last_system(int precision, float* ms)
{
	UT_REQUIRE(has_system, "solver: no system assembled yet (run a step first)");
	UT_REQUIRE(!abd_mode(), "solver: the probe re-solve is for the cloth system (affine bodies solve in q-space)");
	UT_REQUIRE(precision == kPcgFp64 || precision == kPcgMixed, "solver: probe precision must be FP64 or mixed");
	UT_REQUIRE(precision == kPcgFp64 || p.pcg.precision != kPcgFp64,
		"solver: the FP32 matrix copy exists only when the solver runs the mixed or FP32 PCG");
	const hipStream_t s = ctx->stream;
	dx_probe.reserve(3 * (size_t)d.nv);
	PcgParams pp = p.pcg;
	pp.precision = precision;
	UT_HIP_CHECK(hipEventRecord(ev[0], s));
	const PcgReport r = pcg.solve(pp, precision == kPcgFp64 ? (const void*)vals64.get() : (const void*)vals32.get(),
		rhs.get(), dx_probe.get(), precision == kPcgMixed ? vals64.get() : nullptr);
	UT_HIP_CHECK(hipEventRecord(ev[1], s));
	UT_HIP_CHECK(hipEventSynchronize(ev[1]));
	if (ms)
		UT_HIP_CHECK(hipEventElapsedTime(ms, ev[0], ev[1]));
	return r;
}

} // namespace ut
