// Device scratch for ut_graph_components_static (see graph_static.hip).
#pragma once

#include <stdint.h>

namespace ut
{

// Largest graph ut_graph_components_static accepts: the scratch array holds
// 2 parents per vertex (8 MiB per GPU).
constexpr uint32_t kGraphStaticMaxVertices = 1u << 20;

// Looks up the scratch array's address on `device` (which must be current);
// once per device, thread-safe.
void graph_static_setup(int device);
// The address found by graph_static_setup for `device`.
uint32_t* graph_static_scratch(int device);

} // namespace ut
