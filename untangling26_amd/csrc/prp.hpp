// PRP untangling on the GPU: per-contour ray casting of candidate separation
// directions, hit clustering, J(r) evaluation and response pairs.
//
// Replaces host_resolve_intersections_PRP / device_resolve_intersections_PRP
// (Solver/CollisionDetector/intersection_resolver2.cpp:2750-4069,
// intersection_resolver_gpu.cpp:2792) for the default configuration: one
// cloth object, no solid-interior hit links, no direction optimisation.
//
// Work split.  The GPU does everything proportional to the mesh or to the
// hits: normals, per-contour BFS hop and Bellman-Ford rest-geodesic distances,
// candidate ownership, VF/EE/FV rays through the step's LBVHs, hit adjacency
// with union-find, boundary attachment, EF-anchor coverage through a hash set,
// cluster selection, the J(r) reductions, contour adjacency and the response
// records.  The host runs the adaptive k-ring control (a few scalars per
// (contour, direction) per level), the 3x3 eigen-decompositions of the
// direction generator and the contour suppression order.
#pragma once

#include "contact.hpp"
#include "contour.hpp"
#include "lbvh.hpp"
#include "ut_runtime.hpp"

#include <vector>

namespace ut
{

struct PrpHit
{
	uint32_t slot;
	uint32_t type; // 0 VF (vertex, face), 1 EE (edge, edge), 2 FV (face, vertex)
	uint32_t src, dst;
	float	 b0, b1, dist;
	uint32_t root; // union-find parent, then the component root
};
static_assert(sizeof(PrpHit) == 32, "PrpHit layout");

// One (contour, direction) evaluation at hop radii (ks, kd).
struct PrpSlot
{
	uint32_t contour; // index in the batch
	uint32_t combo;
	uint32_t ks, kd;
	float	 dir[3];
	float	 ee_pmin, ee_pmax; // projections of the candidate target edges on dir
	uint32_t pad;
};
static_assert(sizeof(PrpSlot) == 40, "PrpSlot layout");

struct PrpSummary
{
	uint32_t raw_hits, hit_count, covered, ef_count;
	float	 depth, volume, objective, area, disp_cost;
	uint32_t max_hop_src, max_hop_dst, usable;
	uint32_t root; // kept component, kNone when none
	uint32_t pad[3];
};
static_assert(sizeof(PrpSummary) == 64, "PrpSummary layout");

struct PrpParams
{
	uint32_t max_contours = 256;
	uint32_t total_dirs = 10; // 10 = the reference's generator; more appends Fibonacci directions
	uint32_t batch_contours = 32;
	float	 response_depth = 1.0f;
	float	 kappa = 1e8f;
	float	 tau = 0.25f;
};

struct PrpChoice
{
	int32_t	 combo = -1;
	uint32_t ks = 0, kd = 0;
	uint32_t hit_count = 0, covered = 0, ef_count = 0, levels = 0, suppressed = 0;
	float	 dir[3] = { 0, 0, 0 };
	float	 depth = 0, objective = 0, area = 0, disp_cost = 0;
};

// Launch profile of the EE sweep kernel (the PRP's dominant kernel): HIP-event
// time and the counters its algorithmic bytes are computed from.
struct PrpKernelProfile
{
	uint32_t launches = 0;
	float	 ms = 0;
	uint64_t requests = 0, visits = 0, hits = 0;
	uint64_t leaf_tests = 0; // leaves the traversal reached (each checked against the candidate set)
	uint64_t cand_tests = 0; // of which candidates: the edge-edge test runs on these
	uint64_t wave_steps = 0; // traversal loop steps summed over waves (64 lanes each)
};

struct PrpStats
{
	uint32_t contours = 0, levels = 0, slots = 0, hits = 0, responses = 0;
	float	 ms_prepare = 0, ms_levels = 0, ms_final = 0;
	uint64_t rays[3] = { 0, 0, 0 }; // VF, EE, FV ray requests cast
	PrpKernelProfile ee;
};

struct PrpResolver
{
	~PrpResolver();
	Context* ctx = nullptr;
	uint32_t nv = 0, ne = 0, nf = 0;
	// caller's device arrays (static topology and rest data)
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* fae = nullptr;
	const uint32_t* eaf = nullptr;
	const float*	rest_x = nullptr;
	const float*	va = nullptr;
	const float*	ea = nullptr;
	const float*	fa = nullptr;
	const float*	mass = nullptr;
	PrpParams		params;
	uint32_t		prim_bits = 0, slot_bits = 0;

	// adjacency CSRs (ascending neighbours), built at init
	DevBuf<uint32_t> vv_off, vv, ve_off, ve, vf_off, vf;
	std::vector<uint32_t> h_vv_off, h_vv;

	// per-resolve device state
	DevBuf<float>	 normals; // face [3F], edge [3E], vertex [3V]
	DevBuf<uint32_t> b_off, b_verts;		 // boundary verts per (contour, role), global contour index
	DevBuf<uint32_t> flag_off, flags;		 // per vertex sorted flags 2c+role (non-loop pairs)
	DevBuf<uint32_t> cpair_off, cpair_rec;	 // per contour: (edge v0, v1, face v0, v1, v2, mesh idx) words
	DevBuf<uint32_t> hop;					 // [2 * batch][V]
	DevBuf<float>	 geo, curd;				 // [2 * batch][V]
	DevBuf<uint32_t> queue, inq;			 // distance work lists + (hop, geodesic) words [2 * batch][4V], [2 * batch][V]
	DevBuf<uint32_t> adj_blocks;			 // [V][8] (neighbour, rest length bits) pairs
	DevBuf<uint32_t> dist_req;				 // distance launch requests (slot, hop threshold, first)
	DevBuf<uint32_t> dist_state, dist_out;	 // per slot: (hop threshold, radius), (largest hop, ring geodesic bound)
	// rows of the current batch in hop / geo / h_diam (distances may hold more contours)
	const uint32_t*	 hop_b = nullptr;
	const float*	 geo_b = nullptr;
	const uint32_t*	 diam_b = nullptr;
	DevBuf<uint8_t>	 own;					 // [2 * batch][V] ownership (valid where hop <= k)
	DevBuf<uint32_t> kring;					 // [2 * batch] current k of each (contour, role)
	DevBuf<uint32_t> cand_off, cand;		 // per (contour, role): candidate vertex and edge lists
	DevBuf<uint32_t> cand_cnt;				 // [2 * batch][2]
	uint64_t		 ray_count[3] = { 0, 0, 0 }; // ray requests cast since the last resolve began
	PrpKernelProfile ee_prof;
	hipEvent_t		 ee_ev[2] = { nullptr, nullptr };
	DevBuf<uint32_t>		   keep_pos, scan_tmp; // ranks of kept hits / live responses (flags, then scan)
	DevBuf<unsigned long long> tstat; // [3 kinds][4] traversal counters of one evaluation, [12 + kind] wave steps
	DevBuf<PrpSlot>	 slots;
	DevBuf<float>	 slot_tbox;				 // [6 * slots] box of the candidate target edges (EE culling)
	// raw hits of an evaluation: kHitLists sub-lists of hit_cap / kHitLists entries,
	// each with its own counter (list_count[16 k], 64 B apart), so that appends
	// from different waves do not contend for one counter
	DevBuf<PrpHit>	 hits, hits_sorted;
	DevBuf<uint32_t> list_count;
	DevBuf<uint32_t> hit_dst; // target primitive of each sorted hit
	DevBuf<uint32_t> counters;				 // [0] hits, [1] attached roots, [2] final hits, [3] overflow
	DevBuf<uint32_t> sort_k[2], sort_v[2], sort_h;
	DevBuf<uint32_t> slot_range;			 // [2 * slots] sorted hit range per slot
	DevBuf<uint32_t> comp_info;				 // per hit (as root): [attached, count, key lo, key hi, cov]
	DevBuf<uint32_t> roots;					 // attached roots list
	DevBuf<uint64_t> hash;					 // (generation, root, vertex, side) set
	uint32_t		 hash_gen = 255;		 // generation of the next evaluation (255: clear first)
	DevBuf<uint64_t> buckets;				 // hit bucket key -> first sorted index (generation-tagged)
	uint32_t		 bucket_cap = 0, bucket_gen = 255;
	DevBuf<uint64_t> slot_best;				 // [2 * slots]
	DevBuf<PrpSummary> summary;
	DevBuf<uint64_t>   ecand, fcand; // candidate edge / face bitmaps per (contour, role)
	DevBuf<uint8_t>	   eval_part; // per (slot, share) partial sums of the evaluation
	DevBuf<uint8_t>	 adjm;					 // contour adjacency [nc * nc]
	DevBuf<float>	 contour_c;				 // response constraint per contour
	uint32_t		 hit_cap = 0, hash_cap = 0;

	std::vector<float> h_x; // host copy of positions (directions)

	void init(Context* c, uint32_t nv, uint32_t ne, uint32_t nf, const uint32_t* edges, const uint32_t* faces,
		const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, const float* rest_x, const float* vert_area,
		const float* edge_area, const float* face_area, const float* mass, const PrpParams& p);

	// Resolve the contours of `ct` at positions x: choices[nc]; the response
	// records (CollisionPairTemplate) are left in `responses`.  Returns their count.
	uint32_t resolve(const float* x, const ContourBuilder& ct, Lbvh& face_tree, Lbvh& edge_tree,
		std::vector<PrpChoice>& choices, PrpStats* stats);
	DevBuf<ContactRec> responses;

	// Parity hooks.  prepare(): contour tables, normals, and for the batch of
	// contours [c0, c0 + n) the distances and directions.  eval_slots(): one
	// evaluation of the given slots of that batch with hop radii kr[2 lc + role];
	// the sorted hits (with roots) stay in hits_sorted, the kept root per slot in
	// slot_root.  Returns the raw hit count.
	void prepare(const float* x, const ContourBuilder& ct, uint32_t c0, uint32_t n);
	uint32_t eval_slots(const std::vector<PrpSlot>& s, Lbvh& face_tree, Lbvh& edge_tree,
		const std::vector<uint32_t>& kr, std::vector<PrpSummary>& out);

	// host-side contour data of the last prepare
	uint32_t						nc = 0, batch0 = 0, batch_n = 0, total_pairs = 0;
	std::vector<uint32_t>			h_cpair_off, h_cpairs, h_pair_words, h_b_off, h_b_verts, h_diam, h_dist_out;
	std::vector<uint32_t>			h_contour_of_pair, h_mesh_of_pair, h_adj, h_adjn, h_adj_off;
	std::vector<std::vector<float>> h_dirs;
	std::vector<float>				h_va;
	double							global_c[3] = { 0, 0, 0 };
	const float*					x_dev = nullptr;
	uint32_t						last_raw_hits = 0, last_final_hits = 0;

	// more device state
	DevBuf<uint32_t> target_cnt, req_off, slot_root, status, spill;
	DevBuf<uint64_t> comp_key;
	DevBuf<PrpHit>	 final_hits;
	DevBuf<uint32_t> final_contour;
	DevBuf<float>	 contour_dir;
	DevBuf<uint4>	 response_meta; // per response: (source, target, packed region, kind)
	DevBuf<uint8_t>	 suppressed;
	uint32_t		 final_cap = 0;

private:
	void build_contour_tables(const ContourBuilder& ct);
	void distances(uint32_t c0, uint32_t n);
	void dist_reserve(uint32_t n);
	void dist_pass(uint32_t c0, const std::vector<uint32_t>& req, bool mode_b);
	void dist_results(uint32_t n);
	void begin_batch(uint32_t first);
	void directions(uint32_t c0, uint32_t n);
	void ensure_hits(uint32_t cap);
};

} // namespace ut
