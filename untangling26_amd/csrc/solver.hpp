// One implicit Newton step of the untangling simulator on one GPU:
// NewtonSolver::physics_step_GPU (Solver/SimulationSolver/newton_solver.cpp:2482-2767).
#pragma once

#include "contact.hpp"
#include "contour.hpp"
#include "energy.hpp"
#include "lbvh.hpp"
#include "pcg.hpp"
#include "prp.hpp"
#include "regions.hpp"
#include "abd.hpp"

#include <hip/hip_runtime.h>

namespace ut
{

struct SolverDesc
{
	uint32_t nv = 0, ne = 0, nf = 0, nh = 0;
	// topology (device)
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* face_adj_edges = nullptr;
	const uint32_t* edge_adj_faces = nullptr;
	// per-vertex data (device)
	const float*	rest_x = nullptr; // rest positions (PRP rest geodesics); x0 when null
	const float*	mass = nullptr;
	const float*	d_hat = nullptr;
	const float*	offset = nullptr;
	const uint32_t* vprop = nullptr;
	const float*	rest_vert_area = nullptr;
	const float*	rest_edge_area = nullptr;
	const float*	rest_face_area = nullptr;
	// per-vertex friction coefficient (sa_contact_active_verts_friction_coeff); null: no friction
	const float* friction_mu = nullptr;
	// material (device)
	const float*	dm = nullptr;
	const float*	mu_lambda = nullptr;
	const uint32_t* hinges = nullptr;
	const float*	rest_angle = nullptr;
	const float*	bend_stiff_area = nullptr;
	// block pattern (device)
	uint32_t		nslices = 0, nslots = 0;
	const uint32_t* slice_off = nullptr;
	const uint32_t* col = nullptr;
	const uint32_t* diag_slot = nullptr;
	const uint32_t* face_slots = nullptr;
	const uint32_t* hinge_slots = nullptr;
	// affine bodies (nb > 0): every vertex belongs to a body and the step solves in q-space
	AbdDesc		  abd;
	const double* q0 = nullptr; // [12 nb]
};

struct SolverParams
{
	double dt = 1.0 / 60.0;
	double gravity[3] = { 0.0, -9.8, 0.0 };
	double kappa = 1e8;
	int	   energy = kEnergyBarrier;
	double damping = 2.0;
	int	   nonlinear_iters = 1;
	int	   use_self_collision = 1;
	int	   use_ccd = 1;
	int	   use_untangling = 1; // EF detection + contours every iteration
	int	   use_prp = 1;		   // PRP responses for the contours (use_untangling_PRP)
	int	   quasi_static = 0;   // use_quasi_static_mode: x_tilde = x + h^2 g, no inertia gradient
	int	   qs_inertia_hessian = 1; // QuasiStaticInertiaMode::HessianOnly keeps m / h^2 on the diagonal
	// filter_invalid_pairs / filter_broad_phase_pairs (intersection_resolver.cpp:2594-2682)
	// on the proximity and CCD pairs; the reference runs both whenever use_untangling is set
	int		  region_filter = 1;
	// scene_params.h:55; 0 (the CLI's setting) turns zero-distance proximity pairs into EF
	// pairs (intersection_resolver.cpp:404-421) and filters degenerate EF pairs (:841-897)
	int		  ignore_near_zero_dist_pairs = 1;
	PcgParams pcg;
	PrpParams prp;
};

struct StepReport
{
	uint32_t vf_pairs = 0, ee_pairs = 0, ef_pairs = 0, ef_total = 0, contours = 0;
	uint32_t pcg_iterations = 0, pcg_refinements = 0;
	double	 pcg_relative_residual = 0.0;
	int		 pcg_converged = 0;
	float	 toi = 1.0f;
	double	 energy[3] = { 0, 0, 0 };
	float	 ms_detect = 0, ms_contours = 0, ms_assemble = 0, ms_pcg = 0, ms_ccd = 0, ms_total = 0;
	uint32_t prp_responses = 0, prp_contours = 0, prp_levels = 0, prp_hits = 0;
	float	 ms_prp = 0;
	uint32_t prp_slots = 0;
	uint64_t prp_rays[3] = { 0, 0, 0 };
	PrpKernelProfile prp_ee;
};

struct NewtonSolver
{
	Context*	 ctx = nullptr;
	SolverDesc	 d;
	SolverParams p;

	DevBuf<float>  x, v, x_step_start, x_iter_start, x_end;
	DevBuf<double> x_tilde, grad, rhs, dx, vals64, energy;
	DevBuf<float>  vals32;
	DevBuf<float>  toi; // [1]
	DevBuf<float>  rest_copy;
	Lbvh		   face_tree, edge_tree;
	PcgSolver	   pcg;
	ContourBuilder contours;
	PrpResolver	   prp;
	RegionFilter   regions;
	std::vector<PrpChoice> last_choices;

	// contact lists: VF pairs then EE pairs in one array
	uint32_t			cap_contacts = 0, cap_ef = 0;
	DevBuf<ContactRec>	contact_pairs;
	DevBuf<uint32_t>	contact_ids;
	DevBuf<EfPairRec>	ef_pairs;
	DevBuf<uint32_t>	ef_ids;
	DevBuf<uint32_t>	counts; // [0] VF, [1] EE, [2] EF, [3] total contacts, [5] friction pairs
	DevBuf<ContactRec>	fric_pairs; // the first iteration's proximity records (lagged friction)
	// affine bodies: q, its velocity, step / iteration starts, inertia target, gradient, blocks, step
	DevBuf<double>		q, vq, q_step_start, q_iter_start, q_tilde, gq, rhs_q, dq, q_trial, blocks, abd_energy;
	AbdPcg				abd_pcg;
	bool				abd_mode() const { return d.abd.nb > 0; }
	uint32_t			nfric = 0;
	DevBuf<uint32_t>	vert_adj_off, vert_adj_edges; // edges around each vertex (ascending)
	// deterministic assembly: faces and hinges by colour class (no shared vertex
	// within a class), and the canonical order of the proximity records
	DevBuf<uint32_t>	  face_order, hinge_order;
	std::vector<uint32_t> face_color_off, hinge_color_off;
	bool				  colored = false;
	ContactSorter		  csort;
	uint32_t*			h_counts = nullptr;
	hipEvent_t			ev[8] = {};

	void init(Context* c, const SolverDesc& desc, const SolverParams& params, const float* x0, const float* v0);
	~NewtonSolver();
	StepReport physics_step_GPU();
	// Solves the last Newton iteration's system again (same matrix, contacts and
	// right-hand side) at `precision` into a probe vector, leaving the solver's
	// state alone; *ms is the solve's duration (HIP events on the stream).
	PcgReport resolve_last_system(int precision, float* ms);
	// New parameters between steps: validated, and the state a parameter needs is
	// built on demand (the vertex-edge adjacency of the zero-distance EF pairs when
	// ignore_near_zero_dist_pairs is turned off after set-up, as the reference's
	// CLI does: cli.py:93-118 configures the solver after init_solver).
	void set_params(const SolverParams& params);

	uint32_t last_responses = 0; // PRP response records of the last Newton iteration

	DevBuf<double> dx_probe;
	bool		   has_system = false;

private:
	void detect(StepReport& rep);
	void ccd_line_search();
	void build_vert_adj_edges(); // sa_vert_adj_edges_csr (init_mesh_data.cpp), once at set-up
	void build_colors();		 // colour classes of faces and hinges, once at set-up
	void ensure_contact_capacity(uint32_t need_contacts, uint32_t need_ef, uint32_t keep = 0);
};

} // namespace ut
