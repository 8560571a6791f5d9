// LSD radix sort of (u32 key, u32 value) pairs, stable, 8-bit digits.
//
// Replaces the host round trip + CpuParallel::parallel_sort of the reference
// LBVH build (lbvh.cpp:1023-1047).  Keys there are (morton30 << 32 | index),
// unique, so a stable sort of the 30-bit code with the index as value yields
// the same total order.
//
// Per pass three launches:
//   upsweep     per-tile digit histogram in LDS          -> hist[digit][tile]
//   digit scan  one workgroup per digit: exclusive scan over tiles + digit total
//   downsweep   digit base (scan of totals) + stable per-tile ranking with
//               wave64 ballot multisplit, scatter
#include "ut_common.hpp"
#include "ut_runtime.hpp"
#include "radix_sort.hpp"

namespace ut
{

constexpr int kRsThreads = 256;
constexpr int kRsKpt = 8;
constexpr int kRsTile = kRsThreads * kRsKpt;
constexpr int kRsBins = 256;
constexpr int kRsWaves = kRsThreads / kWave;

__global__ __launch_bounds__(kRsThreads) void k_rs_upsweep(
	const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist, uint32_t num_tiles)
{
	__shared__ uint32_t h[kRsBins];
	const uint32_t		tid = threadIdx.x;
	const uint32_t		tile = blockIdx.x;
	h[tid] = 0;
	__syncthreads();
	const uint32_t base = tile * kRsTile;
#pragma unroll
	for (int k = 0; k < kRsKpt; ++k)
	{
		const uint32_t i = base + k * kRsThreads + tid;
		if (i < n)
			atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
	}
	__syncthreads();
	hist[tid * num_tiles + tile] = h[tid];
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts
// (coalesced, 256 at a time), total of the digit into totals[digit].
__global__ __launch_bounds__(256) void k_rs_digit_scan(uint32_t* __restrict__ hist, uint32_t num_tiles,
	uint32_t* __restrict__ totals)
{
	__shared__ uint32_t wsum[256 / kWave];
	const uint32_t		d = blockIdx.x;
	const uint32_t		tid = threadIdx.x;
	uint32_t*			row = hist + (size_t)d * num_tiles;
	uint32_t			carry = 0;
	for (uint32_t base = 0; base < num_tiles; base += 256)
	{
		const uint32_t i = base + tid;
		const uint32_t c = i < num_tiles ? row[i] : 0u;
		uint32_t	   v = c;
		for (int o = 1; o < kWave; o <<= 1)
		{
			const uint32_t t = __shfl_up(v, o, kWave);
			if ((int)__lane_id() >= o)
				v += t;
		}
		const uint32_t w = tid / kWave;
		if (__lane_id() == kWave - 1)
			wsum[w] = v;
		__syncthreads();
		uint32_t before = carry;
		for (uint32_t q = 0; q < w; ++q)
			before += wsum[q];
		uint32_t chunk = 0;
		for (uint32_t q = 0; q < 256 / kWave; ++q)
			chunk += wsum[q];
		if (i < num_tiles)
			row[i] = before + v - c;
		carry += chunk;
		__syncthreads();
	}
	if (tid == 0)
		totals[d] = carry;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_downsweep(const uint32_t* __restrict__ keys_in,
	const uint32_t* __restrict__ vals_in,
	uint32_t* __restrict__ keys_out,
	uint32_t* __restrict__ vals_out,
	uint32_t n,
	int		 shift,
	const uint32_t* __restrict__ hist,
	uint32_t num_tiles,
	const uint32_t* __restrict__ totals)
{
	__shared__ uint32_t wave_cnt[kRsWaves][kRsBins];
	__shared__ uint32_t digit_base[kRsBins];
	__shared__ uint32_t wsum[kRsWaves];
	const uint32_t		tid = threadIdx.x;
	const uint32_t		tile = blockIdx.x;
	const uint32_t		w = tid / kWave;
	{
		// exclusive scan of the 256 digit totals
		const uint32_t c = totals[tid];
		uint32_t	   v = c;
		for (int o = 1; o < kWave; o <<= 1)
		{
			const uint32_t t = __shfl_up(v, o, kWave);
			if ((int)__lane_id() >= o)
				v += t;
		}
		if (__lane_id() == kWave - 1)
			wsum[w] = v;
		__syncthreads();
		uint32_t before = 0;
		for (uint32_t q = 0; q < w; ++q)
			before += wsum[q];
		digit_base[tid] = before + v - c + hist[tid * num_tiles + tile];
	}
#pragma unroll
	for (int q = 0; q < kRsWaves; ++q)
		wave_cnt[q][tid] = 0;

	const uint32_t base = tile * kRsTile;
	uint32_t	   k_reg[kRsKpt], v_reg[kRsKpt];
#pragma unroll
	for (int k = 0; k < kRsKpt; ++k)
	{
		const uint32_t i = base + k * kRsThreads + tid;
		k_reg[k] = i < n ? keys_in[i] : 0u;
		v_reg[k] = i < n ? vals_in[i] : 0u;
	}
	__syncthreads();

	const uint64_t lt = lanemask_lt();
#pragma unroll
	for (int k = 0; k < kRsKpt; ++k)
	{
		const uint32_t i = base + k * kRsThreads + tid;
		const bool	   valid = i < n;
		const uint32_t d = (k_reg[k] >> shift) & 0xFFu;
		uint64_t	   peers = __ballot(valid);
#pragma unroll
		for (int b = 0; b < 8; ++b)
		{
			const bool	   bit = (d >> b) & 1u;
			const uint64_t m = __ballot(bit);
			peers &= bit ? m : ~m;
		}
		const uint32_t rank = (uint32_t)__popcll(peers & lt);
		if (valid && rank == 0)
			wave_cnt[w][d] = (uint32_t)__popcll(peers);
		__syncthreads();
		if (valid)
		{
			uint32_t pos = digit_base[d] + rank;
			for (uint32_t q = 0; q < w; ++q)
				pos += wave_cnt[q][d];
			keys_out[pos] = k_reg[k];
			vals_out[pos] = v_reg[k];
		}
		__syncthreads();
		uint32_t add = 0;
#pragma unroll
		for (int q = 0; q < kRsWaves; ++q)
		{
			add += wave_cnt[q][tid];
			wave_cnt[q][tid] = 0;
		}
		digit_base[tid] += add;
		__syncthreads();
	}
}

void radix_sort_pairs(RadixSortBuffers& b, uint32_t n, int key_bits, hipStream_t stream)
{
	if (n == 0)
		return;
	const uint32_t tiles = div_up(n, kRsTile);
	UT_REQUIRE(b.hist && b.keys[0] && b.keys[1] && b.vals[0] && b.vals[1], "radix sort buffers missing");
	int src = 0;
	for (int shift = 0; shift < key_bits; shift += 8)
	{
		const int dst = src ^ 1;
		k_rs_upsweep<<<tiles, kRsThreads, 0, stream>>>(b.keys[src], n, shift, b.hist, tiles);
		k_rs_digit_scan<<<kRsBins, 256, 0, stream>>>(b.hist, tiles, b.hist + (size_t)tiles * kRsBins);
		k_rs_downsweep<<<tiles, kRsThreads, 0, stream>>>(
			b.keys[src], b.vals[src], b.keys[dst], b.vals[dst], n, shift, b.hist, tiles, b.hist + (size_t)tiles * kRsBins);
		src = dst;
	}
	UT_LAUNCH_CHECK();
	b.result = src;
}

uint32_t radix_sort_hist_size(uint32_t n) { return (div_up(n ? n : 1, kRsTile) + 1) * kRsBins; }

} // namespace ut
