#pragma once

#include "contact_rec.hpp"
#include "region_view.hpp"
#include "ut_runtime.hpp"
#include "ef_dcd.hpp"
#include "lbvh.hpp"

namespace ut
{

enum ContactEnergy : int
{
	kEnergyQuadratic = 0, // contact_energy_type 0
	kEnergyBarrier = 1,	  // contact_energy_type 1 (log barrier, the default)
};

struct ContactArgs
{
	uint32_t		num_verts = 0, num_edges = 0;
	const float*	x = nullptr;	 // DCD positions / CCD start
	const float*	x_end = nullptr; // CCD end positions
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* vprop = nullptr;
	const float*	d_hat = nullptr;  // per vertex
	const float*	offset = nullptr; // per vertex contact offset (thickness)
	const float*	rest_vert_area = nullptr;
	const float*	rest_edge_area = nullptr;
	const float*	rest_face_area = nullptr;
	double			kappa = 1e8;
	int				energy = kEnergyBarrier;
	uint32_t*		count = nullptr;
	ContactRec*		pairs = nullptr;
	uint2*			pair_ids = nullptr;
	uint32_t		max_pairs = 0;
	float*			toi = nullptr; // CCD: global minimum time of impact (non-negative float)
	RegionView		filter;		   // CCD: pairs across a contour or at invalid vertices are skipped
};

// kind: 0 VF DCD, 1 EE DCD, 2 VF CCD, 3 EE CCD.  With `tree` the query is fused
// with the traversal (face tree for VF, edge tree for EE, refit beforehand);
// otherwise it runs over `broad` pairs.
void contact_query(Context* ctx, int kind, const ContactArgs& a, Lbvh* tree, const uint32_t* broad,
	const uint32_t* broad_count, uint32_t broad_cap);
void toi_reset(Context* ctx, float* toi);

// Puts the proximity records [0, n) and their (query, leaf) ids in the order of
// (kind, query, leaf), kind 0 (VF) for [0, nvf) and 1 (EE) after: the detection
// appends with atomics, so its order changes from run to run, while everything
// summed over the records afterwards (assembly, the PCG's contact products)
// must see one fixed order to be bit-reproducible.  query_bits / leaf_bits bound
// the ids.  VF records stay first.
struct ContactSorter
{
	DevBuf<uint32_t>   keys[2], vals[2], hist;
	DevBuf<ContactRec> tmp;
	DevBuf<uint32_t>   tmp_ids;
	void sort(Context* ctx, ContactRec* recs, uint32_t* ids, uint32_t n, uint32_t nvf, int query_bits, int leaf_bits);
};
enum PotentialTerms : int
{
	kTermBarrier = 1,  // repulsion (barrier or quadratic) re-solved at x
	kTermFriction = 2, // lagged friction of the records with a friction multiplier
};

struct PotentialArgs
{
	const ContactRec* pairs = nullptr;
	const uint32_t*	  count = nullptr;
	uint32_t		  cap = 0;
	const float*	  x = nullptr;	 // positions (FP32) ...
	const double*	  x64 = nullptr; // ... or FP64 (then every position array is FP64)
	const float*	  x_start = nullptr;
	const double*	  x_start64 = nullptr; // lagged positions of the friction term
	const float*	  d_hat = nullptr;
	const float*	  offset = nullptr;
	double			  kappa = 1e8;
	int				  energy = kEnergyBarrier;
	int				  terms = kTermBarrier;
	double*			  grad = nullptr; // optional double[3V]: += dE/dx
	double*			  out = nullptr;  // double[2 + 512]: barrier, friction energy, then scratch
};

// Contact potential at x: barrier and / or friction energies, optionally their gradient.
void contact_potential(Context* ctx, const PotentialArgs& a);
// First-iteration friction values of the records (mu_lambda = -mu k1, dv = 0), in place.
void friction_init(Context* ctx, ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* mu);
// Later iterations: relative motion since x_start, repulsion off; out may alias pairs.
void friction_lagged(Context* ctx, const ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const double* x64, const float* x_start, const double* x_start64, ContactRec* out);
// Sum of the contact energies of the records at x (FP64 distances recomputed from
// x); *out is a device double, scratch holds 256 doubles.
void contact_energy(Context* ctx, const ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const float* d_hat, const float* offset, double kappa, int energy, double* out, double* scratch);

} // namespace ut
