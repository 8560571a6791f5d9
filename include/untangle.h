/*
 * untangle.h -- C ABI of libuntangle_hip, the gfx950 implementation of the
 * Untangling26 per-step hot path (LBVH, EF intersection detection, contours,
 * proximity DCD, ACCD, IPC assembly, PCG).
 *
 * Conventions
 *  - Every compute entry point takes DEVICE pointers (HBM) and plain sizes.
 *    Positions are packed float[3*V]; edges uint32[2*E]; faces uint32[3*F].
 *  - Work is enqueued on the context's HIP stream; nothing synchronises unless
 *    the function says so.  Device-side counters may exceed the capacity the
 *    caller passed: the caller compares and retries with a larger buffer,
 *    exactly like the reference's resize-and-retry loop
 *    (newton_solver.cpp:1402-1438).
 *  - Return value: 0 ok, >0 recoverable (1 capacity, 2 not converged),
 *    <0 fatal (-1 invalid argument, -2 HIP error, -3 internal).  The message
 *    of the last failure is returned by ut_last_error(ctx).  No C++ exception
 *    crosses this boundary.
 *  - One stream per context; a context is not thread-safe (the reference
 *    drives one luisa::compute::Stream from one thread, solver_interface.h:39-67).
 *  - A context lives until ut_ctx_destroy has been called AND every object
 *    created from it has been destroyed, so objects may be destroyed after
 *    their context, in any order.
 *
 * Reference citations are paths inside ChengzhuUwU/Untangling26.
 */
#ifndef UNTANGLE_H
#define UNTANGLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UT_API __attribute__((visibility("default")))

typedef struct ut_ctx      ut_ctx;
typedef struct ut_lbvh     ut_lbvh;
typedef struct ut_contours ut_contours;
typedef struct ut_pcg      ut_pcg;
typedef struct ut_solver   ut_solver;

/* ---- context ---------------------------------------------------------- */

/* ABI version of this header. */
UT_API int ut_version(void);
enum { UT_CTX_OWN_STREAM = 1 };
/* Creates a context on HIP device `device`.  Kernels are launched on
 * `hip_stream` (a hipStream_t; NULL is the device's default stream) unless
 * flags has UT_CTX_OWN_STREAM, in which case the context creates and owns a
 * non-blocking stream.  Replaces the luisa::compute::Device/Stream pair of
 * solver_interface.h:35-67. */
UT_API int ut_ctx_create(int device, void* hip_stream, int flags, ut_ctx** out);
UT_API int ut_ctx_destroy(ut_ctx* ctx);
UT_API const char* ut_last_error(const ut_ctx* ctx);
/* Blocks until all work on the context stream is done. */
UT_API int ut_ctx_synchronize(ut_ctx* ctx);
/* The hipStream_t the context launches on. */
UT_API void* ut_ctx_stream(ut_ctx* ctx);

/* ---- LBVH: lcs::LBVH, Solver/CollisionDetector/lbvh.h:164-230 ---------- */

enum { UT_TREE_FACE = 1, UT_TREE_EDGE = 2 };
enum { UT_QUERY_VERTS = 0, UT_QUERY_EDGES = 1, UT_QUERY_EDGES_UNCULLED = 2 };

/* LbvhData::allocate (lbvh.h:108-150) for a face or edge tree of num_leaves. */
UT_API int ut_lbvh_create(ut_ctx* ctx, int tree_type, uint32_t num_leaves, ut_lbvh** out);
UT_API int ut_lbvh_destroy(ut_lbvh* bvh);
/* LBVH::reduce_face_tree_aabb / reduce_edge_tree_aabb (lbvh.cpp:1003-1021)
 * followed by the Morton pass of construct_tree (fn_compute_mortons,
 * lbvh.cpp:512-528).  prims: faces uint32[3N] or edges uint32[2N]. */
UT_API int ut_lbvh_reduce_aabb(ut_lbvh* bvh, const float* x, const uint32_t* prims);
/* LBVH::construct_tree (lbvh.cpp:1023-1047): on-device radix sort instead of
 * the host sort, then fn_apply_sorted + fn_build_inner_nodes. */
UT_API int ut_lbvh_construct_tree(ut_lbvh* bvh);
/* Builds the topology from caller-supplied sorted 64-bit keys
 * (morton30 << 32 | object index), as LBVH::unit_test does (lbvh.cpp:920-984). */
UT_API int ut_lbvh_build_from_sorted_keys(ut_lbvh* bvh, const uint64_t* sorted_keys);
/* LBVH::update_{face,edge}_tree_leave_aabb + LBVH::refit (lbvh.cpp:1050-1092):
 * swept leaf box of x_start/x_end inflated by thickness[first vertex]. */
UT_API int ut_lbvh_refit(ut_lbvh* bvh, const float* thickness, const float* x_start, const float* x_end,
	const uint32_t* prims);
/* LBVH::broad_phase_query_from_{verts,edges,edges_unculled} (lbvh.cpp:1193-1262).
 * Appends (query id, leaf primitive id) pairs to list (uint32[2*max_count]) and
 * adds the number found to *count (device; not reset by this call). */
UT_API int ut_lbvh_query(ut_lbvh* bvh, int kind, const float* x_begin, const float* x_end, const uint32_t* edges,
	uint32_t num_queries, const float* d_hat, const float* thickness, uint32_t* count, uint32_t* list,
	uint32_t max_count);
/* Device-to-device copies of the tree in the reference numbering (inner nodes
 * 0..N-2, leaves N-1..2N-2); any output may be NULL.  children is uint32[2(2N-1)]. */
UT_API int ut_lbvh_tree_arrays(ut_lbvh* bvh, uint64_t* morton_sorted, uint32_t* sorted_get_original,
	uint32_t* parent, uint32_t* children, uint32_t* object_idx);
/* Root box after refit (device float[6]) and global centroid box (device float[6]). */
UT_API int ut_lbvh_boxes(ut_lbvh* bvh, float* root_box, float* global_box);
/* Host-visible health: *healthy = 1 when the build succeeded, *truncated =
 * number of queries whose traversal stack was exhausted (always 0 for trees
 * built from unique keys).  Synchronises the stream. */
UT_API int ut_lbvh_status(ut_lbvh* bvh, uint32_t* healthy, uint32_t* truncated);
/* make_morton32 of lbvh.cpp:109-121 for normalised points (device float[3n]). */
UT_API int ut_morton30(ut_ctx* ctx, const float* points, uint32_t n, uint32_t* codes);

/* ---- EF intersection DCD: IntersectionResolver::ef_dcd_intersection_query
 *      (intersection_resolver.cpp:365-423, kernel :86-197) ---------------- */

/* broad: (eid, fid) pairs from UT_QUERY_EDGES_UNCULLED on the face tree;
 * broad_count: device count of them (pairs beyond broad_cap are ignored).
 * vprop: optional per-vertex property words (bit0 fixed, bit1 rigid,
 * bits 8.. object id).  Emits EfPair records (48 B: 5 x u32 indices, edge
 * bary, face bary x/y, gradient[3], stiffness) and (eid, fid) into pairs /
 * pair_ids, adding the number found to *count (device). */
UT_API int ut_ef_dcd(ut_ctx* ctx, const uint32_t* broad, const uint32_t* broad_count, uint32_t broad_cap,
	const float* x, const uint32_t* edges, const uint32_t* faces, const uint32_t* vprop, const float* d_hat,
	uint32_t* count, void* pairs, uint32_t* pair_ids, uint32_t max_pairs);

/* Fused form of broad_phase_query_from_edges_unculled + fn_narrow_phase_ef_dcd_query
 * (newton_solver.cpp:1283-1297 followed by intersection_resolver.cpp:365-423):
 * traverses the refit face tree with every edge's box and runs the EF test on
 * each candidate in the same kernel, so the broad EF list is never stored.
 * Same output set as the two-step form. */
UT_API int ut_ef_detect(ut_lbvh* face_tree, const float* x, const uint32_t* edges, uint32_t num_edges,
	const uint32_t* faces, const uint32_t* vprop, const float* d_hat, const float* thickness, uint32_t* count,
	void* pairs, uint32_t* pair_ids, uint32_t max_pairs);

/* fn_add_ef_from_proximity_pairs (intersection_resolver.cpp:199-363), run by
 * ef_dcd_intersection_query (:404-421) when ignore_near_zero_dist_pairs is
 * false: every VF / EE proximity pair (recs: CollisionPairTemplate records,
 * rec_ids: (vertex, face) or (edge, edge), *rec_count of them, at most
 * rec_cap read) whose weighted point difference has |d|^2 < 1e-10 appends EF
 * pairs to (pairs, pair_ids), incrementing *count as ut_ef_detect does.
 * vert_adj_off uint32[V+1] / vert_adj_edges: the edges around each vertex
 * (sa_vert_adj_edges_csr); edge_adj_faces as for ut_contours_create. */
UT_API int ut_ef_from_proximity(ut_ctx* ctx, const void* recs, const uint32_t* rec_ids, const uint32_t* rec_count,
	uint32_t rec_cap, const float* x, const uint32_t* edges, const uint32_t* faces, const uint32_t* vert_adj_off,
	const uint32_t* vert_adj_edges, const uint32_t* edge_adj_faces, uint32_t* count, void* pairs,
	uint32_t* pair_ids, uint32_t max_pairs);

/* ---- contours: IntersectionResolver::host_preprocess_ef_pairs
 *      (intersection_resolver.cpp:683-1096) + intersection_contour_construction_from_ext_adjacent
 *      (intersection_resolver2.cpp:254-309, contour_flood_filling_template
 *      intersection_resolver_helper.cpp:518-575) ----------------------------- */

enum {
	UT_CT_PAIR_IDS = 0,		 /* uint32[2*total] (eid, fid), canonical order, loop pairs appended */
	UT_CT_PAIRS = 1,		 /* EfPair[total] (12 words each) */
	UT_CT_CONTOUR_INDEX = 2, /* uint32[total] ef_pair_contour_index */
	UT_CT_MESH_INDEX = 3,	 /* uint32[total] ef_pair_mesh_index (parity) */
	UT_CT_ADJ = 4,			 /* uint32[num_adj] ef_pair_adj_pairs_ext, neighbours packed by pair */
	UT_CT_ADJ_N = 5,		 /* uint32[total] neighbour count of each pair */
	UT_CT_INCONSISTENT = 6,	 /* uint32[total] 1 at the smallest pair of an odd-cycle component */
	UT_CT_COMP_MIN = 7,		 /* uint32[total] smallest pair index of the pair's contour */
	UT_CT_ADJ_OFF = 8,		 /* uint32[total+1] start of each pair's neighbours in UT_CT_ADJ */
	UT_CT_CONTOUR_TYPE = 9,	 /* uint32[contours] classify_contour_type (intersection_resolver2.cpp:311-375):
							    0 undefined, 1 closed, 2 eight, 3 LL, 4 BLI, 5 cross, 6 BIBI, 7 BBII */
	UT_CT_MESH_CLOSED = 10	 /* uint32[2*contours] mesh_is_closed of roles 0 and 1 (:2823-2841) */
};
/* Mesh topology (device arrays kept by reference): face_adj_edges uint32[3F]
 * (edge between face[j] and face[j+1], init_mesh_data.cpp:419-483) and
 * edge_adj_faces uint32[2E] (faces around each edge in ascending order,
 * 0xFFFFFFFF when absent; init_mesh_data.cpp:486-500). */
UT_API int ut_contours_create(ut_ctx* ctx, uint32_t nv, uint32_t ne, uint32_t nf, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, ut_contours** out);
UT_API int ut_contours_destroy(ut_contours* ct);
/* num_pairs EF pairs (EfPair records + (eid, fid)) in any order.  Sorts them
 * canonically ((eid, fid), then payload bytes), drops degenerate pairs with no
 * non-degenerate neighbour when ignore_near_zero_dist_pairs is 0
 * (intersection_resolver.cpp:841-897), adds loop-vertex pairs, builds the EF
 * adjacency and the contours.  Synchronises; *num_total = EF pairs including
 * loop pairs (what narrow_phase_collision_count[1] holds after
 * preprocessing), *num_contours. */
UT_API int ut_contours_build(ut_contours* ct, uint32_t num_pairs, const void* pairs, const uint32_t* pair_ids,
	int ignore_near_zero_dist_pairs, uint32_t* num_total, uint32_t* num_contours);
/* Counts of the last build: out4 = {pairs given, pairs after the degenerate
 * filter, total with loop pairs, adjacency entries}. */
UT_API int ut_contours_counts(ut_contours* ct, uint32_t* out4);
/* Device-to-device copy of one UT_CT_* output of the last build. */
UT_API int ut_contours_get(ut_contours* ct, int what, void* dst);
/* Components with parity of an undirected CSR graph with 0/1 flip weights
 * (UnitTest/test_parallel_contour_construction.cpp:281-448).  scratch is
 * uint32[2n]; outputs uint32[n]. */
UT_API int ut_graph_components(ut_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
	const uint32_t* flip_w, uint32_t* scratch, uint32_t* comp_min, uint32_t* parity, uint32_t* inconsistent);
/* ut_graph_components without the scratch argument, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * union-find parents live in a static device array of 2 * 2^20 uint32 per GPU
 * (8 MiB, held until the process ends); n above 2^20 is rejected (status -1).
 * Calls on one GPU share that array, so calls that overlap on one GPU (other
 * streams, other threads) race; the array keeps the last call's contents and
 * every call re-initialises what it uses.  Same results as ut_graph_components. */
UT_API int ut_graph_components_static(ut_ctx* ctx, uint32_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
	const uint32_t* flip_w, uint32_t* comp_min, uint32_t* parity, uint32_t* inconsistent);

/* ---- PCG: ConjugateGradientSolver (Solver/LinearSolver/precond_cg.h:471-543) ---- */

enum { UT_PCG_FP64 = 0, UT_PCG_MIXED = 1, UT_PCG_FP32 = 2 };
enum { UT_PREC_BLOCK_JACOBI = 0, UT_PREC_POINT_JACOBI = 1 };
/* Krylov recurrence: UT_PCG_CLASSIC is precond_cg.cpp:706-914's loop (two
 * reductions per iteration); UT_PCG_SINGLE_REDUCTION is the Chronopoulos-Gear
 * form of the same method (one reduction point per iteration, two launches
 * instead of three; same iterates in exact arithmetic). */
enum { UT_PCG_CLASSIC = 0, UT_PCG_SINGLE_REDUCTION = 1 };

typedef struct ut_pcg_params {
	int		 precision;		 /* UT_PCG_* */
	int		 preconditioner; /* UT_PREC_* */
	uint32_t max_iters;		 /* total PCG iterations; 0 skips the solve (precond_cg.cpp:443-452) */
	double	 rel_tol;		 /* ||b - A x|| / ||b|| (kPcgRelativeTolerance = 1e-6, precond_cg.cpp:17) */
	double	 inner_tol;		 /* mixed: relative tolerance of each FP32 inner solve */
	uint32_t max_refine;	 /* mixed: FP64 refinement steps */
	double	 regularization; /* lambda: A + lambda * mean(diag block) I (precond_cg.cpp:286-339) */
	uint32_t check_every;	 /* iterations between host convergence polls */
	int		 recurrence;	 /* UT_PCG_CLASSIC or UT_PCG_SINGLE_REDUCTION */
} ut_pcg_params;

/* PcgSolveReport (precond_cg.h:18-28). failure: 0 none, 1 non-finite rhs,
 * 2 non-finite or breakdown during iteration, 3 iteration budget 0. */
typedef struct ut_pcg_report {
	double	 relative_residual; /* final true ||b - A x|| / ||b||, FP64 */
	double	 infinity_norm;		/* max over blocks of |x_i| */
	uint32_t iterations;
	uint32_t refinements;
	int		 converged;
	int		 failure;
} ut_pcg_report;

/* Matrix pattern in SELL-64 block layout (device arrays, kept by reference):
 * slice_off uint32[ceil(n/64)+1], col uint32[nslots], diag_slot uint32[n].
 * Values are planar: vals[c * nslots + slot], c = 3*i + j of the 3x3 block. */
UT_API int ut_pcg_create(ut_ctx* ctx, uint32_t n, uint32_t nslices, uint32_t nslots, const uint32_t* slice_off,
	const uint32_t* col, const uint32_t* diag_slot, ut_pcg** out);
UT_API int ut_pcg_destroy(ut_pcg* pcg);
/* ConjugateGradientSolver::device_solve: solves A x = b (b, x FP64 device
 * arrays of 3n; vals FP64 for UT_PCG_FP64, FP32 otherwise).  Synchronises.
 * Returns 0, or 2 (not converged) with the last iterate in x. */
UT_API int ut_pcg_solve(ut_pcg* pcg, const ut_pcg_params* params, const void* vals, const double* b, double* x,
	ut_pcg_report* report);
/* Mixed-precision solve of the FP64 system: FP32 inner PCG on vals32 (the
 * FP32 rounding of vals64), refinement residuals r = b - A x with vals64, so the
 * converged x solves the FP64 matrix (precond_cg.cpp:706-914 solves in FP32 only;
 * params->precision must be UT_PCG_MIXED). */
UT_API int ut_pcg_solve_refined(ut_pcg* pcg, const ut_pcg_params* params, const float* vals32, const double* vals64,
	const double* b, double* x, ut_pcg_report* report);
/* y = A x (FP64 vectors and accumulation; vals FP32 when vals_fp32 != 0). */
UT_API int ut_bsr_spmv(ut_pcg* pcg, const void* vals, int vals_fp32, const double* x, double* y);
/* Contact pairs (CollisionPairTemplate list, device count) whose off-diagonal
   blocks w_a w_b k2 n n^T are applied matrix-free by ut_pcg_solve* and
   ut_bsr_spmv; their diagonal blocks belong in the block values (ut_assemble
   puts them there).  pairs = NULL detaches.  Replaces the contact triplets the
   reference's SpMV callback sums (newton_solver.cpp:1517-1555). */
UT_API int ut_pcg_set_contacts(ut_pcg* pcg, const void* pairs, const uint32_t* count, uint32_t cap);

/* ---- proximity DCD and CCD: NarrowPhasesDetector (narrow_phase.h:25-346) ---- */

enum { UT_CONTACT_VF_DCD = 0, UT_CONTACT_EE_DCD = 1, UT_CONTACT_VF_CCD = 2, UT_CONTACT_EE_CCD = 3 };
enum { UT_ENERGY_QUADRATIC = 0, UT_ENERGY_BARRIER = 1 };

typedef struct ut_contact_args {
	uint32_t		num_verts, num_edges;
	const float*	x;		/* DCD positions, CCD start positions (float[3V]) */
	const float*	x_end;	/* CCD end positions */
	const uint32_t* edges;
	const uint32_t* faces;
	const uint32_t* vprop;	/* optional per-vertex property words */
	const float*	d_hat;	/* per vertex */
	const float*	offset; /* per vertex contact offset */
	const float*	rest_vert_area;
	const float*	rest_edge_area;
	const float*	rest_face_area;
	double			kappa;	/* stiffness_collision (scene_params.h:116) */
	int				energy; /* UT_ENERGY_* (contact_energy_type) */
	uint32_t*		count;	/* DCD: device count (may exceed max_pairs) */
	void*			pairs;	/* DCD: CollisionPairTemplate[max_pairs] (64 B each) */
	uint32_t*		pair_ids; /* DCD: uint32[2*max_pairs] (vid, fid) or (eid, eid) */
	uint32_t		max_pairs;
	float*			toi;	/* CCD: device float, minimum time of impact (reset with ut_toi_reset) */
} ut_contact_args;

/* kind UT_CONTACT_VF_DCD: fn_narrow_phase_vf_dcd_query (narrow_phase.cpp:980-1157);
 * UT_CONTACT_EE_DCD: fn_narrow_phase_ee_dcd_query (:1159-1346); UT_CONTACT_VF_CCD /
 * UT_CONTACT_EE_CCD: fn_narrow_phase_{vf,ee}_ccd_query (:412-800) with the global
 * TOI minimum (use_global_ccd).  With tree != NULL (the refit face tree for VF,
 * edge tree for EE) the broad phase runs in the same kernel; otherwise the
 * broad pairs of ut_lbvh_query are used. */
UT_API int ut_contact_query(ut_ctx* ctx, int kind, const ut_contact_args* args, ut_lbvh* tree,
	const uint32_t* broad, const uint32_t* broad_count, uint32_t broad_cap);
/* Sets *toi = 1 (line_search_max_t) on the context stream (fn_reset_toi). */
UT_API int ut_toi_reset(ut_ctx* ctx, float* toi);
/* Contact part of the incremental potential at x: for each record the FP64
   closest-point distance d is recomputed from x, then kappa * area * b(d - th,
   d_hat) (UT_ENERGY_BARRIER) or kappa * area * (d - th - d_hat)^2 / 2
   (UT_ENERGY_QUADRATIC), summed over the records with k2 != 0 (disabled pairs
   carry no energy).  The energies whose derivatives narrow_phase.cpp:1091-1112
   assembles; no counterpart entry point in the reference (its line search sums
   them inside the solver).  out: device double[257], out[0] receives the sum,
   out[1..256] is scratch. */
UT_API int ut_contact_energy(ut_ctx* ctx, const void* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const float* d_hat, const float* offset, double kappa, int energy, double* out);

/* Lagged IPC friction of contact records (Friction::ipc_barrier, friction_kernel.hpp:137-301).
 * ut_friction_init is fn_process_collision_pair_friction's first Newton iteration
 * (narrow_phase.cpp:1479-1491): mu_lambda = -mu k1 with mu the mean of the per-vertex
 * coefficients of the record's vertices 0 and 2, zero relative motion; in place on
 * records [0, min(*count, cap)).  ut_friction_lagged is its later iterations
 * (:1492-1506): relative motion sum_k w_k (x_k - x_start_k) with the record's weights,
 * repulsion switched off (disable_repulsion_part); out may equal pairs.  Positions are
 * FP32 (x, x_start) or FP64 (x64, x_start64). */
UT_API int ut_friction_init(ut_ctx* ctx, void* pairs, const uint32_t* count, uint32_t cap, const float* friction_mu);
UT_API int ut_friction_lagged(ut_ctx* ctx, const void* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const double* x64, const float* x_start, const double* x_start64, void* out);

/* Contact part of the incremental potential at x (no reference counterpart: the
 * reference evaluates these terms inside its Newton step, fn_compute_repulsion_energy,
 * narrow_phase.cpp:3432-3530, but exposes no gradient).  UT_TERM_BARRIER: the
 * repulsion of every record with k2 != 0, its closest-point distance re-solved in
 * FP64 at x, and its gradient kappa A b'(d - th) w_k n with the re-solved weights;
 * UT_TERM_FRICTION: the lagged friction of every record with a friction multiplier
 * and its gradient.  out: double[2 + 512] (barrier energy, friction energy, scratch);
 * grad (optional, double[3V]) is added to. */
#define UT_TERM_BARRIER 1
#define UT_TERM_FRICTION 2
typedef struct ut_potential_args {
	const void*		pairs;
	const uint32_t* count;
	uint32_t		cap;
	const float*	x;		   /* FP32 positions [3V], or */
	const double*	x64;	   /* FP64 positions [3V] (then x_start64 as well) */
	const float*	x_start;   /* lagged positions of the friction term */
	const double*	x_start64;
	const float*	d_hat;	/* [V] */
	const float*	offset; /* [V] */
	double			kappa;
	int				energy; /* UT_ENERGY_* */
	int				terms;	/* UT_TERM_* */
	double*			grad;
	double*			out;
} ut_potential_args;
UT_API int ut_contact_potential(ut_ctx* ctx, const ut_potential_args* args);

/* ---- assembly of the incremental potential ------------------------------------------
 * NewtonSolver's material assembly (newton_solver.cpp:2641-2714 with
 * Energies/{soft_inertia_energy,stretch_face_energy,bending_energy_kernel}.cpp)
 * and NarrowPhasesDetector::device_perPair_evaluate_gradient_hessian
 * (narrow_phase.cpp:2619-2746).  Writes grad = dE/dx (double[3V]) and the
 * block values (double[9 * nslots], planar) of the material Hessian plus the
 * contacts' diagonal blocks; vals32 (optional) receives an FP32 copy. */
typedef struct ut_assemble_args {
	uint32_t		nv, nf, nh, nslots;
	const float*	x;
	const double*	x_tilde;	 /* inertia: m / h^2 (x - x_tilde); NULL skips */
	const float*	mass;
	double			inv_h2;
	const uint32_t* faces;
	const float*	dm;
	const float*	face_area;
	const float*	mu_lambda;
	const uint32_t* face_slots;
	const uint32_t* hinges;
	const float*	rest_angle;
	const float*	bend_stiff_area;
	const uint32_t* hinge_slots;
	const void*		contacts;	 /* CollisionPairTemplate list or NULL */
	const uint32_t* contact_count;
	uint32_t		contact_cap;
	const uint32_t* diag_slot;
	double*			grad;
	double*			vals;
	float*			vals32;
	double*			energy;		 /* optional double[4]: inertia, stretch, bending */
	int				quasi_static;		  /* 1: no inertia gradient / energy (soft_inertia_energy.cpp:68-80) */
	int				qs_no_inertia_hessian; /* with quasi_static: drop the m / h^2 diagonal as well */
	const double*	x64;		 /* optional FP64 positions [3V]; when set they replace x */
} ut_assemble_args;
UT_API int ut_assemble(ut_ctx* ctx, const ut_assemble_args* args);

/* PRP parameters (SceneParams PRP fields, scene_params.h:56-80). */
typedef struct ut_prp_params {
	uint32_t max_contours;	 /* untangling_process_contours_count (scene_params.h:67), default 256 */
	uint32_t total_dirs;	 /* directions per contour: 10 = get_raycasting_direction; more append a Fibonacci set */
	uint32_t batch_contours; /* contours evaluated together (PRP_cpu_contour_batch_size, scene_params.h:68) */
	float	 response_depth; /* untangling_response_depth */
	float	 kappa;			 /* stiffness_untangling */
	float	 tau;			 /* PRP_intrinsic_tau */
} ut_prp_params;

/* ---- affine bodies (ABD) ------------------------------------------------------ */

/* Rigid bodies as 12-DOF affine bodies: q_b = (p, a1, a2, a3), x_v = p + [a1 a2 a3] xbar_v
 * (fn_apply_q_to_x_template, newton_solver.cpp:1882-1907).  Device arrays, kept by reference. */
typedef struct ut_abd_desc {
	uint32_t		nb, nv;
	const uint32_t* vert_off; /* [nb + 1] vertices of body b: [vert_off[b], vert_off[b + 1]) */
	const uint32_t* body;	  /* [nv] body of each vertex */
	const float*	xbar;	  /* [3 nv] model position of each vertex */
	const double*	mass4;	  /* [16 nb] compressed 4x4 mass matrix (init_sim_data.cpp:960-1024) */
	const double*	volume;	  /* [nb] rest volume (abd_volume) */
	const double*	kappa;	  /* [nb] orthogonality stiffness (RigidMaterial::stiffness) */
} ut_abd_desc;

/* fn_apply_q_to_x_template: x (FP32 and / or FP64, [3 nv]) from q ([12 nb], FP64). */
UT_API int ut_abd_q_to_x(ut_ctx* ctx, const ut_abd_desc* desc, const double* q, float* x32, double* x64);

/* q-space gradient and 12x12 blocks of AbdInertiaEnergy (abd_inertia_energy.cpp:25-140),
 * AbdOrthoEnergy (abd_ortho_energy.cpp:25-140) and the contacts scattered onto the
 * bodies (add_to_rigid_body, narrow_phase.cpp:2679-2698): grad [12 nb] and blocks
 * [144 nb] (row-major per body, optional) are overwritten; energy [2] (optional):
 * inertia, orthogonality. */
typedef struct ut_abd_assemble_args {
	const double*	q;
	const double*	q_tilde; /* NULL: no inertia */
	double			inv_h2;
	int				quasi_static;		   /* 1: no inertia gradient / energy */
	int				qs_no_inertia_hessian; /* with quasi_static: no inertia Hessian either */
	const void*		contacts;
	const uint32_t* contact_count;
	uint32_t		contact_cap;
	double*			grad;
	double*			blocks;
	double*			energy;
} ut_abd_assemble_args;
UT_API int ut_abd_assemble(ut_ctx* ctx, const ut_abd_desc* desc, const ut_abd_assemble_args* args);
/* gq = J^T gx (accumulate != 0: gq += J^T gx): a vertex-space gradient [3 nv] gathered onto
 * the bodies [12 nb] (no reference counterpart: the differentiable potential's chain
 * rule through x = p + A xbar). */
UT_API int ut_abd_gather(ut_ctx* ctx, const ut_abd_desc* desc, const double* gx, double* gq, int accumulate);
/* y = A p with the blocks and the contacts between bodies (matrix-free), FP64. */
UT_API int ut_abd_apply(ut_ctx* ctx, const ut_abd_desc* desc, const double* blocks, const void* contacts,
	const uint32_t* contact_count, uint32_t contact_cap, const double* p, double* y);
/* Block-Jacobi (12x12) PCG of the q-space system in one workgroup (at most 1600
 * bodies); precision UT_PCG_FP64 or UT_PCG_MIXED (FP32 inner solves, FP64
 * refinement).  Synchronises; returns 0, or 2 when not converged. */
typedef struct ut_abd_pcg ut_abd_pcg;
UT_API int ut_abd_pcg_create(ut_ctx* ctx, uint32_t nb, ut_abd_pcg** out);
UT_API int ut_abd_pcg_destroy(ut_abd_pcg* pcg);
UT_API int ut_abd_pcg_solve(ut_abd_pcg* pcg, const ut_abd_desc* desc, const double* blocks, const void* contacts,
	const uint32_t* contact_count, uint32_t contact_cap, const double* b, double* x, int precision,
	uint32_t max_iters, double rel_tol, ut_pcg_report* report);

/* ---- Newton step: lcs::NewtonSolver (newton_solver.h:283-326) ------------------ */

/* Static mesh data prepared by the initializer (Solver/Initializer/*); all
 * pointers are device arrays kept by reference for the solver's lifetime. */
typedef struct ut_solver_desc {
	uint32_t		nv, ne, nf, nh;
	const uint32_t* edges;			/* [2E] */
	const uint32_t* faces;			/* [3F] */
	const uint32_t* face_adj_edges; /* [3F], see ut_contours_create */
	const uint32_t* edge_adj_faces; /* [2E] */
	const float*	mass;			/* [V] */
	const float*	d_hat;			/* [V] */
	const float*	offset;			/* [V] contact offset */
	const uint32_t* vprop;			/* [V] optional property words */
	const float*	rest_vert_area; /* [V] */
	const float*	rest_edge_area; /* [E] */
	const float*	rest_face_area; /* [F] */
	const float*	dm;				/* [4F] inverse rest 2x2 (d0, d1, d2, d3), Dm = [[d0, d2], [d1, d3]] */
	const float*	mu_lambda;		/* [2F] BW98 stretch and shear stiffness */
	const uint32_t* hinges;			/* [4H] (e0, e1, o0, o1) */
	const float*	rest_angle;		/* [H] */
	const float*	bend_stiff_area; /* [H] */
	uint32_t		nslices, nslots; /* SELL-64 block pattern, see ut_pcg_create */
	const uint32_t* slice_off;
	const uint32_t* col;
	const uint32_t* diag_slot;
	const uint32_t* face_slots;	 /* [9F] slot of block (a, b) of face vertices a, b */
	const uint32_t* hinge_slots; /* [16H] */
	const float*	rest_x;		 /* [3V] rest positions for the PRP rest geodesics; NULL = x0 */
	const float*	friction_mu; /* [V] friction coefficient per vertex (sa_contact_active_verts_friction_coeff,
									the material's friction_mu, default 0.5); NULL: no friction */
	const struct ut_abd_desc* abd; /* affine bodies: every vertex belongs to a body, the step solves in
									  q-space (NULL: cloth, the SELL-64 pattern fields are required) */
	const double*			  q0;  /* [12 nb] initial q of the affine bodies */
} ut_solver_desc;

/* SceneParams fields the step uses (scene_params.h:24-157). */
typedef struct ut_solver_params {
	double		  dt;
	double		  gravity[3];
	double		  kappa;
	int			  energy; /* UT_ENERGY_* */
	double		  damping;
	int			  nonlinear_iters;
	int			  use_self_collision;
	int			  use_ccd;
	int			  use_untangling; /* EF detection + contours each iteration */
	ut_pcg_params pcg;
	int			  use_prp;		/* use_untangling_PRP: PRP responses for the contours */
	int			  quasi_static; /* use_quasi_static_mode with QuasiStaticInertiaMode::HessianOnly */
	ut_prp_params prp;
	int			  region_filter; /* with use_untangling: filter_invalid_pairs on the proximity pairs and
									filter_broad_phase_pairs on the CCD pairs (newton_solver.cpp:1308-1311,
									:2513-2517 run both whenever use_untangling is set) */
	int			  ignore_near_zero_dist_pairs; /* SceneParams default 1; 0 (cli.py:107) adds EF pairs
											  from zero-distance proximity pairs and the degenerate filter */
} ut_solver_params;

typedef struct ut_step_report {
	uint32_t vf_pairs, ee_pairs, ef_pairs, ef_total, contours;
	uint32_t pcg_iterations, pcg_refinements;
	double	 pcg_relative_residual;
	int		 pcg_converged;
	float	 toi;
	double	 energy[3]; /* inertia, stretch, bending at the last assembly */
	float	 ms_detect, ms_contours, ms_assemble, ms_pcg, ms_ccd, ms_total;
	uint32_t prp_responses, prp_contours, prp_levels, prp_hits;
	float	 ms_prp;
	uint32_t prp_slots;
	uint64_t prp_rays_vf, prp_rays_ee, prp_rays_fv;
	/* launch profile of the PRP EE sweep kernel: launches, HIP-event ms, and the
	   requests, node visits, leaf tests and hits its algorithmic bytes count */
	uint32_t prp_ee_launches;
	float	 ms_prp_ee;
	uint64_t prp_ee_requests, prp_ee_visits, prp_ee_leaf_tests, prp_ee_hits;
	uint64_t prp_ee_cand_tests; /* leaves whose target passed the candidate check (edge-edge tests run) */
	uint64_t prp_ee_wave_steps; /* traversal loop steps summed over waves: visits / (64 x steps) is the
								   fraction of lanes that visit a node per step */
} ut_step_report;

UT_API int ut_solver_create(ut_ctx* ctx, const ut_solver_desc* desc, const ut_solver_params* params, const float* x0,
	const float* v0, ut_solver** out);
UT_API int ut_solver_destroy(ut_solver* s);
/* NewtonSolver::physics_step_GPU (newton_solver.cpp:2482-2767).  Synchronises. */
UT_API int ut_solver_step(ut_solver* s, ut_step_report* report);
UT_API int ut_solver_set_params(ut_solver* s, const ut_solver_params* params);
/* Device-to-device copies of the state: float[3V]. */
UT_API int ut_solver_get_state(ut_solver* s, float* x, float* v);
UT_API int ut_solver_set_state(ut_solver* s, const float* x, const float* v);
/* UT_CT_* output of the contours built in the last step (see ut_contours_get). */
UT_API int ut_solver_contour_get(ut_solver* s, int what, void* dst);
/* Measurement hook (no counterpart in the reference): average duration in ms,
 * from HIP events on the solver's stream, of `reps` launches of one kernel of
 * the step on the state of the last step.  UT_PROFILE_PCG_SPMV: the PCG
 * iteration's SELL SpMV on the last assembled matrix (FP32 values unless the
 * PCG runs in FP64). */
enum { UT_PROFILE_PCG_SPMV = 0 };
UT_API int ut_solver_profile(ut_solver* s, int what, uint32_t reps, float* avg_ms);
/* Positions of the solver (device float[3V]): which 0 = current, 1 = start of
 * the last Newton iteration (the positions the last detection, contours and
 * PRP saw; sa_x_iter_start), 2 = start of the last step (sa_x_step_start). */
UT_API int ut_solver_get_positions(ut_solver* s, int which, float* x);
/* Affine bodies: q and its velocity (FP64 [12 nb] device arrays; either may be NULL).
 * ut_solver_set_q also recomputes the vertex positions x = x(q). */
UT_API int ut_solver_get_q(ut_solver* s, double* q, double* vq);
UT_API int ut_solver_set_q(ut_solver* s, const double* q, const double* vq);
/* The PRP response pairs of the last Newton iteration (target_point_template_pairs
 * and target_point_template_pairs_indices, intersection_resolver2.cpp:2105-2221):
 * *count = their number; the first min(count, cap) records (64 B each) and index
 * quadruples (source, target, packed region 2 * 256 * contour, kind 0 VF / 1 FV /
 * 2 EE) are copied to the device arrays records and meta. */
UT_API int ut_solver_responses(ut_solver* s, void* records, uint32_t* meta, uint32_t cap, uint32_t* count);
/* Measurement hook (no counterpart in the reference): solves the last Newton
 * iteration's system again -- the same assembled matrix, contact blocks and
 * right-hand side -- at `precision` (UT_PCG_FP64, or UT_PCG_MIXED when the
 * solver itself runs mixed) into a probe vector, leaving the solver's state
 * unchanged.  *ms (may be null) receives the solve's duration from HIP events
 * on the solver's stream.  Returns 0, or 2 when the solve did not reach the
 * tolerance.  Lets a caller compare the two precisions on identical systems. */
UT_API int ut_solver_pcg_resolve(ut_solver* s, int precision, ut_pcg_report* report, float* ms);

/* ---- PRP untangling: IntersectionResolver::host_resolve with PRP
 *      (intersection_resolver.cpp:2131-2217 -> intersection_resolver2.cpp:2750-4069) ---- */

typedef struct ut_prp ut_prp;


/* PRPMinContourHitInfo essentials of the chosen direction of one contour. */
typedef struct ut_prp_choice {
	int32_t	 combo; /* direction slot, -1 when the contour has no response */
	uint32_t ks, kd, hit_count, covered, ef_count, levels, suppressed;
	float	 dir[3], depth, objective, area, disp_cost;
	uint32_t pad;
} ut_prp_choice;

typedef struct ut_prp_stats {
	uint32_t contours, levels, slots, hits, responses;
	float	 ms_prepare, ms_levels, ms_final;
	uint64_t rays_vf, rays_ee, rays_fv; /* ray requests cast, per kind */
} ut_prp_stats;

/* One (contour, direction) evaluation at hop radii (ks, kd); contour is the
 * index inside the batch given to ut_prp_prepare. */
typedef struct ut_prp_slot {
	uint32_t contour, combo, ks, kd;
	float	 dir[3], ee_pmin, ee_pmax;
	uint32_t pad;
} ut_prp_slot;

typedef struct ut_prp_summary {
	uint32_t raw_hits, hit_count, covered, ef_count;
	float	 depth, volume, objective, area, disp_cost;
	uint32_t max_hop_src, max_hop_dst, usable, root, pad[3];
} ut_prp_summary;

/* RayCasting::HitInfo (intersection_resolver_helper.h:260-448) compacted:
 * type 0 VF (src vertex, dst face), 1 EE (src edge, dst edge), 2 FV (src face,
 * dst vertex); b0, b1 the face or edge barycentres; root its cluster. */
typedef struct ut_prp_hit {
	uint32_t slot, type, src, dst;
	float	 b0, b1, dist;
	uint32_t root;
} ut_prp_hit;

/* Topology and rest data are the caller's device arrays (kept by reference). */
UT_API int ut_prp_create(ut_ctx* ctx, uint32_t nv, uint32_t ne, uint32_t nf, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, const float* rest_x,
	const float* vert_area, const float* edge_area, const float* face_area, const float* mass,
	const ut_prp_params* params, ut_prp** out);
UT_API int ut_prp_destroy(ut_prp* p);
UT_API int ut_prp_set_params(ut_prp* p, const ut_prp_params* params);
/* host_resolve_intersections_PRP: contours of ct at positions x (device), rays
 * through the LBVHs built at x; the response pairs (CollisionPairTemplate, 64 B)
 * are copied to out_pairs (device, out_cap records) and *num_pairs receives
 * their count; returns 1 (capacity) when out_cap was too small.  choices[c] for
 * every contour (choices_cap entries).  Synchronises. */
UT_API int ut_prp_resolve(ut_prp* p, const float* x, ut_contours* ct, ut_lbvh* face_tree, ut_lbvh* edge_tree,
	void* out_pairs, uint32_t out_cap, uint32_t* num_pairs, ut_prp_choice* choices, uint32_t choices_cap,
	ut_prp_stats* stats);
/* Parity hooks: prepare contour batch [c0, c0 + n) (distances, directions),
 * read its directions and distances, evaluate slots and read the sorted hits of
 * the last evaluation, read the face/edge/vertex normals.  ut_prp_prepare
 * computes the full hop and rest-geodesic fields (and the hop diameters) that
 * ut_prp_distances returns; ut_prp_resolve searches each field only as far as
 * its k-rings reach (exact up to those limits), so read distances after a
 * prepare, not after a resolve. */
UT_API int ut_prp_prepare(ut_prp* p, const float* x, ut_contours* ct, uint32_t c0, uint32_t n);
UT_API int ut_prp_directions(ut_prp* p, uint32_t contour, float* dirs, uint32_t* count);
UT_API int ut_prp_distances(ut_prp* p, uint32_t local_contour, uint32_t role, uint32_t* hop, float* geo,
	uint32_t* diameter);
UT_API int ut_prp_eval(ut_prp* p, ut_lbvh* face_tree, ut_lbvh* edge_tree, const ut_prp_slot* slots, uint32_t nslots,
	const uint32_t* kring, ut_prp_summary* out, uint32_t* raw_hits);
UT_API int ut_prp_hits(ut_prp* p, ut_prp_hit* out, uint32_t cap);
UT_API int ut_prp_normals(ut_prp* p, float* out);

#ifdef __cplusplus
}
#endif

#endif /* UNTANGLE_H */
