#pragma once
#include "ut_runtime.hpp"
#include "lbvh.hpp"

namespace ut
{

// Per-vertex property word (VertexProperty, collision_data.h): bit 0 fixed,
// bit 1 rigid body, bits 8.. object id.
constexpr uint32_t kPropFixed = 1u;
constexpr uint32_t kPropRigid = 2u;
constexpr int	   kPropObjectShift = 8;

// CollisionPair::EfPair (SimulationCore/collision_data.h:152-220): 48 B.
struct EfPairRec
{
	uint32_t idx[5];  // edge v0, v1, face v0, v1, v2
	float	 bary[3]; // edge bary (1 - t), face bary x, face bary y
	float	 grad[3];
	float	 stiff;
};
static_assert(sizeof(EfPairRec) == 48, "EfPair layout");

struct EfDcdArgs
{
	const uint32_t* broad = nullptr;	   // (eid, fid) pairs
	const uint32_t* broad_count = nullptr; // device count of broad pairs
	uint32_t		broad_cap = 0;		   // pairs stored in `broad`
	const float*	x = nullptr;
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* vprop = nullptr; // optional
	const float*	d_hat = nullptr;
	uint32_t*		count = nullptr;	// device: number of EF pairs (may exceed max_pairs)
	uint32_t*		pairs = nullptr;	// EfPairRec[max_pairs]
	uint32_t*		pair_ids = nullptr; // uint2[max_pairs]: (eid, fid)
	uint32_t		max_pairs = 0;
};

// fn_add_ef_from_proximity_pairs (intersection_resolver.cpp:199-363): EF pairs
// made from the zero-distance VF / EE proximity pairs, appended to the EF list.
struct EfFromProximityArgs
{
	const uint32_t* recs = nullptr;		 // CollisionPairTemplate[rec_cap], 16 words each
	const uint32_t* rec_ids = nullptr;	 // uint2[rec_cap]: (vertex, face) or (edge, edge)
	const uint32_t* rec_count = nullptr; // device count of proximity pairs (clamped to rec_cap)
	uint32_t		rec_cap = 0;
	const float*	x = nullptr;
	const uint32_t* edges = nullptr;
	const uint32_t* faces = nullptr;
	const uint32_t* vert_adj_off = nullptr;	  // [V + 1]
	const uint32_t* vert_adj_edges = nullptr; // edges around each vertex
	const uint32_t* edge_adj_faces = nullptr; // [2E], kNone when absent
	uint32_t*		count = nullptr;		  // device EF count, incremented
	uint32_t*		pairs = nullptr;		  // EfPairRec[max_pairs]
	uint32_t*		pair_ids = nullptr;		  // uint2[max_pairs]
	uint32_t		max_pairs = 0;
};

void ef_dcd(Context* ctx, const EfDcdArgs& a);
void ef_from_proximity(Context* ctx, const EfFromProximityArgs& a);
// Broad (edge query on the refit face tree, unculled) + narrow in one kernel;
// a.broad / a.broad_count / a.broad_cap are unused.
void ef_detect(Context* ctx, Lbvh& face_tree, const EfDcdArgs& a, uint32_t num_edges, const float* thickness);

} // namespace ut
