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

void ef_dcd(Context* ctx, const EfDcdArgs& a);
// Broad (edge query on the refit face tree, unculled) + narrow in one kernel;
// a.broad / a.broad_count / a.broad_cap are unused.
void ef_detect(Context* ctx, Lbvh& face_tree, const EfDcdArgs& a, uint32_t num_edges, const float* thickness);

} // namespace ut
