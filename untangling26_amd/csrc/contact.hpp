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
// Sum of the contact energies of the records at x (FP64 distances recomputed from
// x); *out is a device double, scratch holds 256 doubles.
void contact_energy(Context* ctx, const ContactRec* pairs, const uint32_t* count, uint32_t cap, const float* x,
	const float* d_hat, const float* offset, double kappa, int energy, double* out, double* scratch);

} // namespace ut
