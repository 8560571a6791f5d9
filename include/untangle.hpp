/*
 * untangle.hpp -- C++ facade of libuntangle_hip keeping the class and method
 * names of the reference's Solver/ API, header-only over the C ABI of
 * untangle.h:
 *
 *   lcs::LBVH                    Solver/CollisionDetector/lbvh.h:164-230
 *   lcs::NarrowPhasesDetector    Solver/CollisionDetector/narrow_phase.h:25-346
 *   lcs::IntersectionResolver    Solver/CollisionDetector/intersection_resolver.h:23-206
 *   lcs::ConjugateGradientSolver Solver/LinearSolver/precond_cg.h:471-543
 *   lcs::NewtonSolver            Solver/SimulationSolver/newton_solver.h:283-326
 *
 * The reference's methods take a luisa::compute::Stream&; here every object
 * belongs to an lcs::Device (a HIP device and the stream its kernels run on).
 * Buffers are the caller's device pointers, as with the C ABI.  A failed call
 * throws lcs::Error carrying the status code and ut_last_error's message, as
 * the reference's C++ code throws std::runtime_error.  Objects own their
 * handles (move-only).
 */
#ifndef UNTANGLE_HPP
#define UNTANGLE_HPP

#include "untangle.h"

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace lcs
{

struct Error : std::runtime_error
{
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace detail
{
inline int check(int rc, const ut_ctx* ctx, const char* what)
{
	if (rc < 0)
		throw Error(rc, std::string(what) + ": " + (ctx ? ut_last_error(ctx) : "no context"));
	return rc;
}

template <typename H, int (*Destroy)(H*)>
class Handle
{
public:
	Handle() = default;
	explicit Handle(H* h) : h_(h) {}
	Handle(const Handle&) = delete;
	Handle& operator=(const Handle&) = delete;
	Handle(Handle&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
	Handle& operator=(Handle&& o) noexcept
	{
		reset();
		h_ = std::exchange(o.h_, nullptr);
		return *this;
	}
	~Handle() { reset(); }
	void reset()
	{
		if (h_)
			(void)Destroy(h_);
		h_ = nullptr;
	}
	H* get() const { return h_; }

private:
	H* h_ = nullptr;
};
} // namespace detail

/* A HIP device and its stream (the reference's luisa::compute::Device + Stream,
 * solver_interface.h:35-67).  stream = nullptr launches on the device's default
 * stream; own_stream creates a non-blocking stream owned by the device object. */
class Device
{
public:
	explicit Device(int device = 0, void* stream = nullptr, bool own_stream = false)
	{
		ut_ctx* h = nullptr;
		const int rc = ut_ctx_create(device, stream, own_stream ? UT_CTX_OWN_STREAM : 0, &h);
		if (rc < 0)
			throw Error(rc, "ut_ctx_create: no HIP device " + std::to_string(device) + " or stream");
		ctx_ = detail::Handle<ut_ctx, ut_ctx_destroy>(h);
	}
	ut_ctx* get() const { return ctx_.get(); }
	void*	stream() const { return ut_ctx_stream(ctx_.get()); }
	void	synchronize() const { detail::check(ut_ctx_synchronize(ctx_.get()), ctx_.get(), "synchronize"); }

private:
	detail::Handle<ut_ctx, ut_ctx_destroy> ctx_;
};

enum class TreeType : int
{
	Face = UT_TREE_FACE,
	Edge = UT_TREE_EDGE,
};

/* lcs::LBVH: a face or edge tree of num_leaves primitives. */
class LBVH
{
public:
	LBVH(Device& dev, TreeType type, uint32_t num_leaves) : dev_(&dev), type_(type)
	{
		ut_lbvh* h = nullptr;
		detail::check(ut_lbvh_create(dev.get(), (int)type, num_leaves, &h), dev.get(), "ut_lbvh_create");
		h_ = detail::Handle<ut_lbvh, ut_lbvh_destroy>(h);
	}
	ut_lbvh* get() const { return h_.get(); }

	void reduce_face_tree_aabb(const float* x, const uint32_t* faces) { reduce(x, faces); }
	void reduce_edge_tree_aabb(const float* x, const uint32_t* edges) { reduce(x, edges); }
	void construct_tree() { detail::check(ut_lbvh_construct_tree(get()), dev_->get(), "construct_tree"); }
	/* The leaf boxes are updated and refit together (update_*_leave_aabb + refit). */
	void update_face_tree_leave_aabb(const float* thickness, const float* x_start, const float* x_end,
		const uint32_t* faces)
	{
		leaves_ = { thickness, x_start, x_end, faces };
	}
	void update_edge_tree_leave_aabb(const float* thickness, const float* x_start, const float* x_end,
		const uint32_t* edges)
	{
		leaves_ = { thickness, x_start, x_end, edges };
	}
	void refit()
	{
		if (!leaves_.x_start)
			throw Error(-1, "LBVH::refit: update_*_tree_leave_aabb first");
		detail::check(ut_lbvh_refit(get(), leaves_.thickness, leaves_.x_start, leaves_.x_end, leaves_.prims),
			dev_->get(), "refit");
	}
	void broad_phase_query_from_verts(const float* x_begin, const float* x_end, uint32_t num_verts,
		const float* d_hat, const float* thickness, uint32_t* count, uint32_t* list, uint32_t max_count)
	{
		query(UT_QUERY_VERTS, x_begin, x_end, nullptr, num_verts, d_hat, thickness, count, list, max_count);
	}
	void broad_phase_query_from_edges(const float* x_begin, const float* x_end, const uint32_t* edges,
		uint32_t num_edges, const float* d_hat, const float* thickness, uint32_t* count, uint32_t* list,
		uint32_t max_count)
	{
		query(UT_QUERY_EDGES, x_begin, x_end, edges, num_edges, d_hat, thickness, count, list, max_count);
	}
	void broad_phase_query_from_edges_unculled(const float* x_begin, const float* x_end, const uint32_t* edges,
		uint32_t num_edges, const float* d_hat, const float* thickness, uint32_t* count, uint32_t* list,
		uint32_t max_count)
	{
		query(UT_QUERY_EDGES_UNCULLED, x_begin, x_end, edges, num_edges, d_hat, thickness, count, list, max_count);
	}

private:
	void reduce(const float* x, const uint32_t* prims)
	{
		detail::check(ut_lbvh_reduce_aabb(get(), x, prims), dev_->get(), "reduce_tree_aabb");
	}
	void query(int kind, const float* xb, const float* xe, const uint32_t* edges, uint32_t n, const float* d_hat,
		const float* th, uint32_t* count, uint32_t* list, uint32_t max_count)
	{
		detail::check(ut_lbvh_query(get(), kind, xb, xe, edges, n, d_hat, th, count, list, max_count), dev_->get(),
			"broad_phase_query");
	}
	struct Leaves
	{
		const float *	thickness = nullptr, *x_start = nullptr, *x_end = nullptr;
		const uint32_t* prims = nullptr;
	};
	Device*									 dev_;
	TreeType								 type_;
	Leaves									 leaves_;
	detail::Handle<ut_lbvh, ut_lbvh_destroy> h_;
};

/* Per-vertex and topology arrays the narrow phase reads (device pointers). */
struct ContactMesh
{
	uint32_t		num_verts = 0, num_edges = 0;
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* vprop = nullptr;
	const float*	d_hat = nullptr;
	const float*	offset = nullptr;
	const float*	rest_vert_area = nullptr;
	const float*	rest_edge_area = nullptr;
	const float*	rest_face_area = nullptr;
	double			kappa = 1e8;
	int				energy = UT_ENERGY_BARRIER;
};

/* lcs::NarrowPhasesDetector: proximity DCD records and the CCD time of impact. */
class NarrowPhasesDetector
{
public:
	NarrowPhasesDetector(Device& dev, const ContactMesh& mesh) : dev_(&dev), m_(mesh) {}

	/* fn_narrow_phase_{vf,ee}_dcd_query with the fused broad phase of the refit tree;
	 * appends CollisionPairTemplate records (64 B) and ids, *count on the device. */
	void vf_dcd_query_repulsion(const float* x, uint32_t* count, void* pairs, uint32_t* pair_ids, uint32_t max_pairs,
		LBVH& face_tree)
	{
		run(UT_CONTACT_VF_DCD, dcd(x, count, pairs, pair_ids, max_pairs), face_tree);
	}
	void ee_dcd_query_repulsion(const float* x, uint32_t* count, void* pairs, uint32_t* pair_ids, uint32_t max_pairs,
		LBVH& edge_tree)
	{
		run(UT_CONTACT_EE_DCD, dcd(x, count, pairs, pair_ids, max_pairs), edge_tree);
	}
	void reset_toi(float* toi) { detail::check(ut_toi_reset(dev_->get(), toi), dev_->get(), "reset_toi"); }
	/* fn_narrow_phase_{vf,ee}_ccd_query: global minimum time of impact in *toi. */
	void vf_ccd_query(const float* x_start, const float* x_end, float* toi, LBVH& face_tree)
	{
		run(UT_CONTACT_VF_CCD, ccd(x_start, x_end, toi), face_tree);
	}
	void ee_ccd_query(const float* x_start, const float* x_end, float* toi, LBVH& edge_tree)
	{
		run(UT_CONTACT_EE_CCD, ccd(x_start, x_end, toi), edge_tree);
	}

private:
	ut_contact_args base() const
	{
		ut_contact_args a{};
		a.num_verts = m_.num_verts, a.num_edges = m_.num_edges, a.edges = m_.edges, a.faces = m_.faces;
		a.vprop = m_.vprop, a.d_hat = m_.d_hat, a.offset = m_.offset;
		a.rest_vert_area = m_.rest_vert_area, a.rest_edge_area = m_.rest_edge_area;
		a.rest_face_area = m_.rest_face_area, a.kappa = m_.kappa, a.energy = m_.energy;
		return a;
	}
	ut_contact_args dcd(const float* x, uint32_t* count, void* pairs, uint32_t* ids, uint32_t max_pairs) const
	{
		ut_contact_args a = base();
		a.x = x, a.count = count, a.pairs = pairs, a.pair_ids = ids, a.max_pairs = max_pairs;
		return a;
	}
	ut_contact_args ccd(const float* xs, const float* xe, float* toi) const
	{
		ut_contact_args a = base();
		a.x = xs, a.x_end = xe, a.toi = toi;
		return a;
	}
	void run(int kind, const ut_contact_args& a, LBVH& tree)
	{
		detail::check(ut_contact_query(dev_->get(), kind, &a, tree.get(), nullptr, nullptr, 0), dev_->get(),
			"contact query");
	}
	Device*		dev_;
	ContactMesh m_;
};

/* Mesh data of the intersection resolver (device pointers, kept by reference). */
struct ResolverMesh
{
	uint32_t		num_verts = 0, num_edges = 0, num_faces = 0;
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* face_adj_edges = nullptr;
	const uint32_t* edge_adj_faces = nullptr;
	const float*	rest_x = nullptr;
	const float*	rest_vert_area = nullptr;
	const float*	rest_edge_area = nullptr;
	const float*	rest_face_area = nullptr;
	const float*	mass = nullptr;
};

/* lcs::IntersectionResolver: EF intersection DCD, the preprocessed EF list and
 * its contours, and the PRP resolve that turns them into response pairs. */
class IntersectionResolver
{
public:
	IntersectionResolver(Device& dev, const ResolverMesh& mesh, const ut_prp_params& params) : dev_(&dev), m_(mesh)
	{
		ut_contours* c = nullptr;
		detail::check(ut_contours_create(dev.get(), mesh.num_verts, mesh.num_edges, mesh.num_faces, mesh.edges,
						  mesh.faces, mesh.face_adj_edges, mesh.edge_adj_faces, &c),
			dev.get(), "ut_contours_create");
		contours_ = detail::Handle<ut_contours, ut_contours_destroy>(c);
		ut_prp* p = nullptr;
		detail::check(ut_prp_create(dev.get(), mesh.num_verts, mesh.num_edges, mesh.num_faces, mesh.edges,
						  mesh.faces, mesh.face_adj_edges, mesh.edge_adj_faces, mesh.rest_x, mesh.rest_vert_area,
						  mesh.rest_edge_area, mesh.rest_face_area, mesh.mass, &params, &p),
			dev.get(), "ut_prp_create");
		prp_ = detail::Handle<ut_prp, ut_prp_destroy>(p);
	}

	/* ef_dcd_intersection_query (intersection_resolver.cpp:365-423), broad phase fused
	 * with the refit face tree: EfPair records (48 B) and (eid, fid) ids. */
	void ef_dcd_intersection_query(LBVH& face_tree, const float* x, const float* d_hat, const float* thickness,
		uint32_t* count, void* pairs, uint32_t* pair_ids, uint32_t max_pairs, const uint32_t* vprop = nullptr)
	{
		detail::check(ut_ef_detect(face_tree.get(), x, m_.edges, m_.num_edges, m_.faces, vprop, d_hat, thickness,
						  count, pairs, pair_ids, max_pairs),
			dev_->get(), "ef_dcd_intersection_query");
	}
	/* host_preprocess_ef_pairs + intersection_contour_construction_from_ext_adjacent
	 * (intersection_resolver.cpp:683-1096, intersection_resolver2.cpp:254-375) on the
	 * device; returns (EF pairs with loop pairs, contours). */
	std::pair<uint32_t, uint32_t> host_preprocess_ef_pairs(uint32_t num_pairs, const void* pairs,
		const uint32_t* pair_ids, bool ignore_near_zero_dist_pairs = true)
	{
		uint32_t total = 0, nc = 0;
		detail::check(ut_contours_build(contours_.get(), num_pairs, pairs, pair_ids, ignore_near_zero_dist_pairs ? 1 : 0,
						  &total, &nc),
			dev_->get(), "host_preprocess_ef_pairs");
		num_contours_ = nc;
		return { total, nc };
	}
	/* One UT_CT_* output of the last preprocessing (device destination). */
	void get(int what, void* dst) { detail::check(ut_contours_get(contours_.get(), what, dst), dev_->get(), "get"); }
	/* host_resolve with PRP: response CollisionPairTemplate records into out_pairs;
	 * returns their number (may exceed out_cap: call again with a larger buffer). */
	uint32_t host_resolve(const float* x, LBVH& face_tree, LBVH& edge_tree, void* out_pairs, uint32_t out_cap,
		std::vector<ut_prp_choice>* choices = nullptr, ut_prp_stats* stats = nullptr)
	{
		uint32_t n = 0;
		ut_prp_choice* ch = nullptr;
		uint32_t	   cap = 0;
		if (choices)
		{
			choices->assign(num_contours_ ? num_contours_ : 1, ut_prp_choice{});
			ch = choices->data(), cap = (uint32_t)choices->size();
		}
		detail::check(ut_prp_resolve(prp_.get(), x, contours_.get(), face_tree.get(), edge_tree.get(), out_pairs,
						  out_cap, &n, ch, cap, stats),
			dev_->get(), "host_resolve");
		return n;
	}
	ut_contours* contours() const { return contours_.get(); }
	ut_prp*		 prp() const { return prp_.get(); }

private:
	Device*											 dev_;
	ResolverMesh									 m_;
	uint32_t										 num_contours_ = 0;
	detail::Handle<ut_contours, ut_contours_destroy> contours_;
	detail::Handle<ut_prp, ut_prp_destroy>			 prp_;
};

/* lcs::ConjugateGradientSolver on a SELL-64 3x3 block pattern. */
class ConjugateGradientSolver
{
public:
	ConjugateGradientSolver(Device& dev, uint32_t n, uint32_t nslices, uint32_t nslots, const uint32_t* slice_off,
		const uint32_t* col, const uint32_t* diag_slot)
		: dev_(&dev)
	{
		ut_pcg* h = nullptr;
		detail::check(ut_pcg_create(dev.get(), n, nslices, nslots, slice_off, col, diag_slot, &h), dev.get(),
			"ut_pcg_create");
		h_ = detail::Handle<ut_pcg, ut_pcg_destroy>(h);
	}
	/* device_solve (precond_cg.cpp:706-914): FP64 or FP32 values; returns the report
	 * (last_solve_report); a solve that stops at its budget is reported, not thrown. */
	const ut_pcg_report& device_solve(const ut_pcg_params& params, const void* vals, const double* b, double* x)
	{
		detail::check(ut_pcg_solve(h_.get(), &params, vals, b, x, &last_), dev_->get(), "device_solve");
		return last_;
	}
	/* FP32 inner solves with FP64 refinement of the FP64 system. */
	const ut_pcg_report& device_solve_refined(const ut_pcg_params& params, const float* vals32, const double* vals64,
		const double* b, double* x)
	{
		detail::check(ut_pcg_solve_refined(h_.get(), &params, vals32, vals64, b, x, &last_), dev_->get(),
			"device_solve_refined");
		return last_;
	}
	void spmv(const void* vals, bool vals_fp32, const double* x, double* y)
	{
		detail::check(ut_bsr_spmv(h_.get(), vals, vals_fp32 ? 1 : 0, x, y), dev_->get(), "spmv");
	}
	void set_contacts(const void* pairs, const uint32_t* count, uint32_t cap)
	{
		detail::check(ut_pcg_set_contacts(h_.get(), pairs, count, cap), dev_->get(), "set_contacts");
	}
	const ut_pcg_report& last_solve_report() const { return last_; }

private:
	Device*								   dev_;
	detail::Handle<ut_pcg, ut_pcg_destroy> h_;
	ut_pcg_report						   last_{};
};

/* lcs::NewtonSolver: the whole untangling step on the device. */
class NewtonSolver
{
public:
	NewtonSolver(Device& dev, const ut_solver_desc& desc, const ut_solver_params& params, const float* x0,
		const float* v0 = nullptr)
		: dev_(&dev)
	{
		init_solver(desc, params, x0, v0);
	}
	void init_solver(const ut_solver_desc& desc, const ut_solver_params& params, const float* x0, const float* v0)
	{
		ut_solver* h = nullptr;
		detail::check(ut_solver_create(dev_->get(), &desc, &params, x0, v0, &h), dev_->get(), "init_solver");
		h_ = detail::Handle<ut_solver, ut_solver_destroy>(h);
	}
	/* physics_step_GPU (newton_solver.cpp:2482-2767). */
	ut_step_report physics_step_GPU()
	{
		ut_step_report r{};
		detail::check(ut_solver_step(h_.get(), &r), dev_->get(), "physics_step_GPU");
		return r;
	}
	/* The reference's CPU step; this library runs the same step on the GPU. */
	ut_step_report physics_step_CPU() { return physics_step_GPU(); }
	void		   set_params(const ut_solver_params& p)
	{
		detail::check(ut_solver_set_params(h_.get(), &p), dev_->get(), "set_params");
	}
	void get_state(float* x, float* v)
	{
		detail::check(ut_solver_get_state(h_.get(), x, v), dev_->get(), "get_state");
	}
	void set_state(const float* x, const float* v)
	{
		detail::check(ut_solver_set_state(h_.get(), x, v), dev_->get(), "set_state");
	}
	void contour_output(int what, void* dst)
	{
		detail::check(ut_solver_contour_get(h_.get(), what, dst), dev_->get(), "contour_output");
	}
	ut_solver* get() const { return h_.get(); }

private:
	Device*										 dev_;
	detail::Handle<ut_solver, ut_solver_destroy> h_;
};

} // namespace lcs

#endif
