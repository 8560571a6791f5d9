// LSD radix sort of (u32 key, u32 value) pairs, stable, 8-bit digits.
//
// Replaces the host round trip + CpuParallel::parallel_sort of the reference
// LBVH build (lbvh.cpp:1023-1047).  Keys there are (morton30 << 32 | index),
// unique, so a stable sort of the 30-bit code with the index as value yields
// the same total order.
//
// Per pass three launches:
//   upsweep   per-tile digit histogram in LDS            -> hist[digit][tile]
//   scan      one workgroup, exclusive scan over hist     (digit-major)
//   downsweep stable per-tile ranking with wave64 ballot multisplit, scatter
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

// Exclusive scan of `total` counters with one 1024-thread workgroup.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ hist, uint32_t total)
{
	__shared__ uint32_t wsum[1024 / kWave];
	const uint32_t		tid = threadIdx.x;
	const uint32_t		per = (total + 1023) / 1024;
	const uint32_t		beg = tid * per;
	const uint32_t		end = min(beg + per, total);
	uint32_t			local = 0;
	for (uint32_t i = beg; i < end; ++i)
		local += hist[i];
	// inclusive wave scan
	uint32_t v = local;
	for (int o = 1; o < kWave; o <<= 1)
	{
		uint32_t t = __shfl_up(v, o, kWave);
		if ((int)__lane_id() >= o)
			v += t;
	}
	const uint32_t w = tid / kWave;
	if (__lane_id() == kWave - 1)
		wsum[w] = v;
	__syncthreads();
	if (tid < kWave)
	{
		uint32_t s = tid < 1024 / kWave ? wsum[tid] : 0u;
		for (int o = 1; o < kWave; o <<= 1)
		{
			uint32_t t = __shfl_up(s, o, kWave);
			if ((int)__lane_id() >= o)
				s += t;
		}
		if (tid < 1024 / kWave)
			wsum[tid] = s; // inclusive
	}
	__syncthreads();
	uint32_t run = (w ? wsum[w - 1] : 0u) + v - local; // exclusive start of this thread's chunk
	for (uint32_t i = beg; i < end; ++i)
	{
		uint32_t c = hist[i];
		hist[i] = run;
		run += c;
	}
}

__global__ __launch_bounds__(kRsThreads) void k_rs_downsweep(const uint32_t* __restrict__ keys_in,
	const uint32_t* __restrict__ vals_in,
	uint32_t* __restrict__ keys_out,
	uint32_t* __restrict__ vals_out,
	uint32_t n,
	int		 shift,
	const uint32_t* __restrict__ hist,
	uint32_t num_tiles)
{
	__shared__ uint32_t wave_cnt[kRsWaves][kRsBins];
	__shared__ uint32_t digit_base[kRsBins];
	const uint32_t		tid = threadIdx.x;
	const uint32_t		tile = blockIdx.x;
	const uint32_t		w = tid / kWave;
	digit_base[tid] = hist[tid * num_tiles + tile];
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
		k_rs_scan<<<1, 1024, 0, stream>>>(b.hist, tiles * kRsBins);
		k_rs_downsweep<<<tiles, kRsThreads, 0, stream>>>(
			b.keys[src], b.vals[src], b.keys[dst], b.vals[dst], n, shift, b.hist, tiles);
		src = dst;
	}
	UT_LAUNCH_CHECK();
	b.result = src;
}

uint32_t radix_sort_hist_size(uint32_t n) { return div_up(n ? n : 1, kRsTile) * kRsBins; }

} // namespace ut
