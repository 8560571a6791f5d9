// The 64-byte contact record (CollisionPair::CollisionPairTemplate,
// collision_data.h:48-150), shared by the device kernels and host checks.
#pragma once

#include <stdint.h>

namespace ut
{

// CollisionPair::CollisionPairTemplate (collision_data.h:48-150): 64 B.
struct ContactRec
{
	uint32_t idx[4];	// VF: v, f0, f1 | type bit, f2;  EE: a0, a1, b0, b1 | type bit
	float	 normal[3]; // vec1.xyz
	float	 area;		// vec1.w
	float	 k1, k2;	// vec2.xy
	float	 bary[2];	// vec2.zw: VF face bary x, y;  EE edge-a bary x, edge-b bary x
	float	 dv[3];		// vec3.xyz (friction)
	float	 lambda;	// vec3.w
};
static_assert(sizeof(ContactRec) == 64, "CollisionPairTemplate layout");

// Bit 31 of idx[2] marks a VF record and bit 31 of idx[3] an EE record
// (collision_data.h:36-80); every index is masked before it addresses a vertex.
constexpr uint32_t kTypeBit = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7FFFFFFFu;

} // namespace ut
