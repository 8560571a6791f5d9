// Exclusive prefix sum of u32 arrays (reduce-then-scan, three launches):
//   k_scan_local  per 1024-element tile: wave64 shuffle scan + LDS, tile sum
//   k_scan_tiles  one workgroup scans the tile sums (up to 4096 tiles)
//   k_scan_add    adds each tile's offset
#include "ut_common.hpp"
#include "scan.hpp"

namespace ut
{

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanThreads * kScanPerThread;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total)
{
	const uint32_t tid = threadIdx.x, w = tid / kWave;
	uint32_t	   inc = v;
	for (int o = 1; o < kWave; o <<= 1)
	{
		const uint32_t t = __shfl_up(inc, o, kWave);
		if ((int)__lane_id() >= o)
			inc += t;
	}
	if (__lane_id() == kWave - 1)
		wsum[w] = inc;
	__syncthreads();
	uint32_t before = 0, all = 0;
	for (uint32_t q = 0; q < blockDim.x / kWave; ++q)
	{
		if (q < w)
			before += wsum[q];
		all += wsum[q];
	}
	*total = all;
	__syncthreads();
	return before + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(const uint32_t* __restrict__ in,
	uint32_t* __restrict__ out, uint32_t n, uint32_t* __restrict__ tile_sums)
{
	__shared__ uint32_t wsum[kScanThreads / kWave];
	const uint32_t		base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
	uint32_t			v[kScanPerThread];
	uint32_t			s = 0;
#pragma unroll
	for (int k = 0; k < kScanPerThread; ++k)
	{
		v[k] = (base + k < n) ? in[base + k] : 0u;
		s += v[k];
	}
	uint32_t	   total;
	const uint32_t ex = block_exclusive_scan(s, wsum, &total);
	uint32_t	   run = ex;
#pragma unroll
	for (int k = 0; k < kScanPerThread; ++k)
	{
		if (base + k < n)
			out[base + k] = run;
		run += v[k];
	}
	if (threadIdx.x == 0)
		tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t* __restrict__ tile_sums, uint32_t ntiles,
	uint32_t* __restrict__ total_out)
{
	__shared__ uint32_t wsum[1024 / kWave];
	uint32_t			carry = 0;
	for (uint32_t b = 0; b < ntiles; b += 1024)
	{
		const uint32_t i = b + threadIdx.x;
		const uint32_t v = i < ntiles ? tile_sums[i] : 0u;
		uint32_t	   total;
		const uint32_t ex = block_exclusive_scan(v, wsum, &total);
		if (i < ntiles)
			tile_sums[i] = carry + ex;
		carry += total;
	}
	if (threadIdx.x == 0 && total_out)
		*total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* __restrict__ out, uint32_t n,
	const uint32_t* __restrict__ tile_sums)
{
	const uint32_t add = tile_sums[blockIdx.x];
	const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
#pragma unroll
	for (int k = 0; k < kScanPerThread; ++k)
		if (base + k < n)
			out[base + k] += add;
}

uint32_t scan_tmp_size(uint32_t n) { return div_up(n ? n : 1, kScanTile); }

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, uint32_t* total, hipStream_t s)
{
	if (n == 0)
		return;
	const uint32_t tiles = div_up(n, kScanTile);
	k_scan_local<<<tiles, kScanThreads, 0, s>>>(in, out, n, tmp);
	k_scan_tiles<<<1, 1024, 0, s>>>(tmp, tiles, total);
	k_scan_add<<<tiles, kScanThreads, 0, s>>>(out, n, tmp);
}

} // namespace ut
