/*
 * oracle.h -- CPU restatement of the Untangling26 hot path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the product (untangling26_amd/) links or calls it.
 *
 * Built by oracle/Makefile with gcc -O2 -ffp-contract=off: FP32 expressions
 * are evaluated in the same order as the GPU kernels, with IEEE rounding and
 * no fused multiply-add, so geometric predicates give identical bits.
 *
 * Each function cites the reference file:line it restates (paths inside
 * ChengzhuUwU/Untangling26).  Pinning of each function is recorded in the
 * header comment of its definition in the .c file.
 */
#ifndef UT_ORACLE_H
#define UT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Threads of the parallel loops (PCG products and dots, CCD minimum, PRP distance
 * searches and direction sets); 1 (the default) runs everything serially, which
 * the tests use.  Only the CPU baseline timing raises it. */
extern int or_nthreads;
void	   or_set_threads(int n);

/* lbvh.cpp:84-121 make_morton32 on normalised points. */
void or_morton30(const float* p, uint32_t n, uint32_t* out);

/* Full LBVH topology: global box of leaf centroids (lbvh.cpp:400-496),
 * Morton keys (lbvh.cpp:512-528), stable sort by 64-bit key (lbvh.cpp:1032-1040),
 * Karras inner nodes (lbvh.cpp:206-293, 530-600).
 * tree_type 1 = faces (prims uint32[3n]), 2 = edges (uint32[2n]).
 * Outputs sized n / 2n-1 / 2(2n-1). */
void or_lbvh_build(int tree_type, const float* x, const uint32_t* prims, uint32_t n, uint64_t* morton_sorted,
	uint32_t* sorted_get_original, uint32_t* parent, uint32_t* children, uint32_t* object_idx, float* global_box);

/* Karras topology from given sorted unique keys (lbvh.cpp:530-600). */
void or_karras(const uint64_t* keys, uint32_t n, uint32_t* parent, uint32_t* children, uint32_t* object_idx);

/* Box of every leaf after update_*_leave_aabb (lbvh.cpp:621-688): out float[6n]. */
void or_leaf_boxes(int tree_type, const float* xs, const float* xe, const uint32_t* prims, uint32_t n,
	const float* thickness, float* out);

/* Query boxes (lbvh.cpp:780-894): kind 0 vertices (n = #verts), 1/2 edges. */
void or_query_boxes(int kind, const float* xb, const float* xe, const uint32_t* edges, uint32_t n, const float* d_hat,
	const float* thickness, float* out);

/* All (query, leaf) pairs with strict box overlap (aabb.h:94-98).  cull != 0
 * keeps only query < leaf (fn_query_from_edges_v2).  Uses a uniform grid; the
 * result is sorted lexicographically.  Returns the number of pairs; writes at
 * most cap of them to out (uint32[2*cap]). */
size_t or_overlap_pairs(const float* qbox, uint32_t nq, const float* lbox, uint32_t nl, int cull, uint32_t* out,
	size_t cap);
/* Same set by O(nq*nl) enumeration, for pinning or_overlap_pairs. */
size_t or_overlap_pairs_brute(const float* qbox, uint32_t nq, const float* lbox, uint32_t nl, int cull, uint32_t* out,
	size_t cap);

/* LineIntersection (intersection_resolver_helper.cpp:333-481): returns 1 on hit. */
int or_line_intersection(const float* s, const float* e, const float* A, const float* B, const float* C, float* bary,
	float* t);

/* fn_narrow_phase_ef_dcd_query (intersection_resolver.cpp:86-197) over given
 * broad pairs; writes EfPair records (12 words) and (eid, fid); returns count. */
size_t or_ef_dcd(const uint32_t* broad, size_t nb, const float* x, const uint32_t* edges, const uint32_t* faces,
	const uint32_t* vprop, const float* d_hat, uint32_t* pairs, uint32_t* pair_ids, size_t cap);

/* Complete EF detection: edge query boxes vs face leaf boxes (unculled,
 * xb = xe = x) followed by or_ef_dcd; (eid, fid) sorted.  Returns count. */
size_t or_ef_detect(const float* x, uint32_t nv, const uint32_t* edges, uint32_t ne, const uint32_t* faces, uint32_t nf,
	const uint32_t* vprop, const float* d_hat, const float* thickness, uint32_t* pairs, uint32_t* pair_ids, size_t cap);

/* Node boxes of a Karras tree from leaf boxes (indexed by primitive id):
 * node_box float[6(2n-1)], reference numbering. */
void or_lbvh_node_boxes(uint32_t n, const uint32_t* children, const uint32_t* sorted_get_original,
	const float* leaf_box, float* node_box);
/* Stack traversal of that tree (query_template2).  Unsorted output; returns count. */
size_t or_lbvh_query(uint32_t n, const uint32_t* children, const uint32_t* object_idx, const float* node_box,
	const float* qbox, uint32_t nq, int cull, uint32_t* out, size_t cap);

/* init_mesh_data.cpp:419-500: face_adj_edges uint32[3F], edge_adj_faces uint32[2E]. */
int or_mesh_topology(uint32_t ne, uint32_t nf, const uint32_t* edges, const uint32_t* faces, uint32_t* face_adj_edges,
	uint32_t* edge_adj_faces);
/* host_preprocess_ef_pairs (intersection_resolver.cpp:683-1096): canonical
 * order, optional degenerate filter, loop-vertex pairs, packed adjacency
 * (*adj malloc'ed, release with or_free).  Returns the total pair count. */
long or_ef_preprocess(uint32_t P, const uint32_t* in_ids, const uint32_t* in_pairs, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, int filter_degenerate,
	uint32_t* out_ids, uint32_t* out_pairs, uint32_t** adj, uint32_t* adj_n, uint32_t* kept);
void or_free(void* p);
/* is_degenerate_ef_pair (intersection_resolver.cpp:561-569) of one EfPair (12 words). */
int or_ef_degenerate(const uint32_t* pair);
/* fn_add_ef_from_proximity_pairs (intersection_resolver.cpp:199-363): EF pairs
 * from zero-distance VF / EE proximity pairs.  Returns the number generated. */
size_t or_ef_from_proximity(size_t n, const uint32_t* recs, const uint32_t* ids, const float* x, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* vae_off, const uint32_t* vae, const uint32_t* edge_adj_faces,
	uint32_t* out_pairs, uint32_t* out_ids, size_t cap);
/* contour_flood_filling_template (intersection_resolver_helper.cpp:518-575), packed adjacency. */
uint32_t or_contours_dfs(uint32_t N, const uint32_t* pairs, const uint32_t* adj, const uint32_t* adj_n,
	uint32_t* contour_index, uint32_t* mesh_index);
/* cpu_reference of UnitTest/test_parallel_contour_construction.cpp:155-198. */
int or_graph_bfs(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t* comp_min,
	uint32_t* parity, uint32_t* comp_inconsistent);

/* host_solve_core (precond_cg.cpp:367-693) in FP64 on a block-CSR matrix. */
uint32_t or_pcg(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* bl, const double* b, double* x,
	int precond, uint32_t max_iters, double rel_tol, double* out);

/* distance.hpp:69-160, 255-310 (FP64) */
void or_pt_coeff(const double* p, const double* t0, const double* t1, const double* t2, double* out);
void or_ee_coeff(const double* a0, const double* a1, const double* b0, const double* b1, double* out);
/* cipc_kernel.hpp:15-36: deriv 0 = b, 1 = b', 2 = b'' */
double or_barrier(double d, double dhat, int deriv);
/* narrow_phase.cpp:980-1346 pair tests; rec = 16 words of CollisionPairTemplate */
int or_vf_dcd(uint32_t vid, uint32_t fid, const float* x, const uint32_t* faces, const uint32_t* vprop,
	const float* dh, const float* off, const float* rva, const float* rfa, double kappa, int energy, uint32_t* rec);
int or_ee_dcd(uint32_t left, uint32_t right, const float* x, const uint32_t* edges, const uint32_t* vprop,
	const float* dh, const float* off, const float* rea, double kappa, int energy, uint32_t* rec);
/* accd.hpp:48-110 on 4 points x0[12] with displacement dx[12] */
double or_accd(int is_ee, const double* x0, const double* dx, double offset);
double or_toi_postprocess(double toi);
double or_vf_ccd(uint32_t vid, uint32_t fid, const float* xs, const float* xe, const uint32_t* faces,
	const uint32_t* vprop, const float* off);
double or_ee_ccd(uint32_t left, uint32_t right, const float* xs, const float* xe, const uint32_t* edges,
	const uint32_t* vprop, const float* off);
size_t or_dcd_list(int kind, const uint32_t* broad, size_t nb, const float* x, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* vprop, const float* dh, const float* off, const float* rva,
	const float* rea, const float* rfa, double kappa, int energy, uint32_t* recs, uint32_t* ids, size_t cap);
double or_ccd_list(int kind, const uint32_t* broad, size_t nb, const float* xs, const float* xe,
	const uint32_t* edges, const uint32_t* faces, const uint32_t* vprop, const float* off);

/* BW98 face term (stretch_energy.h:200-520): x[9] (3 vertices), dm[4]; grad[9], hess[81]. */
double or_bw98_face(const double* x, const float* dm, double area, double mu, double lam, double* grad, double* hess);
/* dihedral bending hinge (bending_energy.h:164-260): x[12] = (e0, e1, o0, o1); grad[12], hess[144]. */
double or_bending_hinge(const double* x, double theta0, double k, double* grad, double* hess, double* theta_out);

/* Full incremental-potential assembly into the SELL slot numbering (planar vals). */
void or_assemble(uint32_t nv, uint32_t nf, uint32_t nh, const float* x, const double* x_tilde, const float* mass,
	double inv_h2, const uint32_t* faces, const float* dm, const float* area, const float* mu_lambda,
	const uint32_t* face_slots, const uint32_t* hinges, const float* rest_angle, const float* bend,
	const uint32_t* hinge_slots, const uint32_t* contacts, uint32_t ncontacts, const uint32_t* diag_slot,
	uint32_t nslots, double* grad, double* vals, double* energy);
/* or_pcg with matrix-free contact off-diagonal blocks (16-word records). */
uint32_t or_pcg_contacts(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* bl, const uint32_t* contacts,
	uint32_t ncontacts, const double* b, double* x, int precond, uint32_t max_iters, double rel_tol, double* out);

void or_classify_contours(uint32_t N, const uint32_t* pairs, const uint8_t* is_boundary, const uint32_t* adj_n,
	const uint32_t* contour_index, const uint32_t* mesh_index, uint32_t nc, uint32_t* type, uint32_t* closed);

/* PRP untangling path (oracle_prp.c).  Opaque mesh/contour state. */
typedef struct or_prp_mesh or_prp_mesh;
or_prp_mesh* or_prp_mesh_create(uint32_t nv, uint32_t ne, uint32_t nf, const float* x, const float* x_rest,
	const uint32_t* edges, const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces,
	const float* vert_area, const float* edge_area, const float* face_area, const float* mass);
void	 or_prp_mesh_destroy(or_prp_mesh* m);
void	 or_prp_normals(const or_prp_mesh* m, float* fn, float* en, float* vn);
void	 or_prp_distances(const or_prp_mesh* m, uint32_t nb, const uint32_t* bverts, uint32_t* hop, float* geo);
int		 or_prp_owns(float own, float opp, float tau);
int		 or_prp_owns_dual(float rest_own, float rest_opp, float cur_own, float cur_opp, float tau, int use_rest,
			 int use_cur);
uint32_t or_prp_directions(uint32_t np, const float* points, const float* weights, uint32_t nb0, const uint32_t* b0,
	uint32_t nb1, const uint32_t* b1, const float* x, const float* vert_area, uint32_t nv, const double* global_c,
	uint32_t total, float* out);
void	 or_prp_prepare(or_prp_mesh* m, uint32_t N, const uint32_t* pairs, const uint32_t* adj, const uint32_t* adjn,
		 const uint32_t* contour_index, const uint32_t* mesh_index, uint32_t nc, uint32_t total_dirs);
void	 or_prp_prepare_n(or_prp_mesh* m, uint32_t N, const uint32_t* pairs, const uint32_t* adj, const uint32_t* adjn,
		 const uint32_t* contour_index, const uint32_t* mesh_index, uint32_t nc, uint32_t total_dirs,
		 uint32_t n_prepare);
void	 or_prp_choose(or_prp_mesh* m, uint32_t c, void* out);
uint32_t or_prp_eval_slot(or_prp_mesh* m, uint32_t c, const float* dir, uint32_t k_src, uint32_t k_dst,
	void* out_hits, uint32_t cap, void* out_metrics);
void	 or_prp_get_state(const or_prp_mesh* m, uint32_t c, float* dirs, uint32_t* ndirs, uint32_t* diam, uint32_t* nb);
void	 or_prp_set_dirs(or_prp_mesh* m, uint32_t c, uint32_t n, const float* dirs);
const uint32_t* or_prp_hop(const or_prp_mesh* m, uint32_t c, uint32_t role);
const float*	or_prp_geo(const or_prp_mesh* m, uint32_t c, uint32_t role);
int		 or_prp_select(uint32_t n, const float* objective, const uint32_t* covered, const uint32_t* ef_count,
		 const float* depth, const float* cost, const uint32_t* hits, const uint32_t* eligible);
uint32_t or_prp_resolve(or_prp_mesh* m, uint32_t max_contours, float response_depth, float kappa,
	void* choices_out, uint32_t* recs, uint32_t cap);
/* Known-answer access to single steps of the PRP selection (see oracle_prp.c):
 * the k-ring driver with a caller-supplied evaluator, the contour suppression
 * and the penetration summary of a hit list. */
typedef void (*or_prp_eval_fn)(void* user, uint32_t combo, uint32_t ks, uint32_t kd, void* metrics);
int		 or_prp_drive(uint32_t nd, uint32_t ks0, uint32_t kd0, uint32_t Ds, uint32_t Dd, or_prp_eval_fn eval, void* user,
		 void* metrics_out, uint32_t* flags, uint32_t* radii, uint32_t* levels, uint32_t* no_response);
uint32_t or_prp_suppress(uint32_t nc, const uint8_t* adjm, const float* area, uint32_t nactive, const uint32_t* active,
	uint8_t* sup);
void	 or_prp_penetration(uint32_t n, const void* hits, const float* face_area, const float* edge_area,
		 const float* vert_area, const float* mass, const float* face_normal, const float* edge_normal,
		 const float* vert_normal, const uint32_t* edges, const uint32_t* faces, const float* dir, void* out_metrics);
int		 or_prp_hit_attached(const or_prp_mesh* m, uint32_t c, const void* hit);
/* Components of a hit list: bucketed search (brute = 0) or all pairs (brute != 0);
 * root[i] = smallest hit index of hit i's component. */
void or_prp_hit_components(const or_prp_mesh* m, const void* hits, uint32_t n, int brute, uint32_t* root);
uint32_t or_prp_vertex_flags(const or_prp_mesh* m, uint32_t v, uint32_t* out, uint32_t cap);
int		 or_prp_pick_component(uint32_t ncomp, const uint8_t* attached, const uint32_t* cov, const uint32_t* cnt,
		 const uint64_t* key);
/* or_prp_resolve plus meta[2i] = contour and meta[2i + 1] = hit type (0 VF, 1 EE,
 * 2 FV) of response i: the region semantics of host_apply_untangling_constraint
 * (intersection_resolver2.cpp:2339-2345) need the kind, which the VF-layout FV
 * record does not carry. */
uint32_t or_prp_resolve_meta(or_prp_mesh* m, uint32_t max_contours, float response_depth, float kappa,
	void* choices_out, uint32_t* recs, uint32_t* meta, uint32_t cap);

/* ---- lagged IPC friction and the contact potential (oracle_friction.c) ---- */
/* friction_kernel.hpp:137-301: energy; grad[3] and hess[9] (optional) with respect to
 * the relative displacement rel of a pair with normal n and multiplier mu_lambda. */
double or_friction_eval(double mu_lambda, const double* n, const double* rel, double E, double* grad, double* hess);
/* CollisionPairTemplate weights (collision_data.h:131-136) of a 16-word record. */
void or_record_weights(const uint32_t* rec, double* w);
/* The record's 3x3 block k2 n n^T + friction Hessian (K[9], optional) and its point
 * gradient k1 n + friction gradient (g[3], optional). */
void or_pair_block(const uint32_t* rec, double* K, double* g);
/* narrow_phase.cpp:1479-1491: mu_lambda = -mu k1, dv = 0 (in place). */
void or_friction_init(uint32_t* recs, uint32_t n, const float* mu);
/* narrow_phase.cpp:1492-1506: dv = sum w (x - x_start), k1 = k2 = 0 (in place). */
void or_friction_lagged(uint32_t* recs, uint32_t n, const double* x, const double* x_start);
/* Barrier (terms & 1) and friction (terms & 2) energies out[2] at FP64 x; grad[3V]
 * (optional) is added to. */
void or_contact_potential(const uint32_t* recs, uint32_t n, const double* x, const double* x_start, const float* d_hat,
	const float* offset, double kappa, int energy, int terms, double* grad, double* out);

#ifdef __cplusplus
}
#endif
#endif
