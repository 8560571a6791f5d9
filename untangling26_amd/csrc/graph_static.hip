// Union-find scratch of ut_graph_components_static: the parity-doubled
// parent array lives in a file-scope device array instead of a caller
// workspace or a library allocation, for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// Each GPU holds its own copy of `scratch` until the process ends.  In host
// code the array's name is not a device address: its address on a device is
// looked up once per device with hipGetSymbolAddress (graph_static_setup, run
// by ut_ctx_create with that device current) and only that address is handed
// to kernels.  Every call on one GPU shares the array, so calls that overlap
// on one GPU race; the kernels re-initialise every parent they use.
#include "graph_static.hpp"

#include "ut_runtime.hpp"

#include <mutex>

static __device__ __align__(256) uint32_t scratch[2 * (size_t)ut::kGraphStaticMaxVertices];

namespace ut
{

namespace
{
constexpr int kMaxDevices = 64;
std::mutex	  g_mu;
uint32_t*	  g_scratch[kMaxDevices] = {};
} // namespace

void graph_static_setup(int device)
{
	UT_REQUIRE(device >= 0 && device < kMaxDevices, "graph scratch: device index out of range");
	std::lock_guard<std::mutex> lock(g_mu);
	if (g_scratch[device])
		return;
	int cur = -1;
	UT_HIP_CHECK(hipGetDevice(&cur));
	UT_REQUIRE(cur == device, "graph scratch: set-up must run with its device current");
	void* p = nullptr;
	UT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
	UT_REQUIRE(p != nullptr, "graph scratch: no device address for the scratch array");
	g_scratch[device] = static_cast<uint32_t*>(p);
}

uint32_t* graph_static_scratch(int device)
{
	UT_REQUIRE(device >= 0 && device < kMaxDevices, "graph scratch: device index out of range");
	std::lock_guard<std::mutex> lock(g_mu);
	UT_REQUIRE(g_scratch[device] != nullptr, "graph scratch: not set up on this device (create a context first)");
	return g_scratch[device];
}

} // namespace ut
