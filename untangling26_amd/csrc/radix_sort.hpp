#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut
{
// Ping-pong storage for radix_sort_pairs.  After the sort the sorted pairs are
// in keys[result] / vals[result].
struct RadixSortBuffers
{
	uint32_t* keys[2] = { nullptr, nullptr };
	uint32_t* vals[2] = { nullptr, nullptr };
	uint32_t* hist = nullptr; // radix_sort_hist_size(n) entries
	int		  result = 0;
};

uint32_t radix_sort_hist_size(uint32_t n);
// Stable ascending sort of the low `key_bits` bits of keys[0] with vals[0] as payload.
void radix_sort_pairs(RadixSortBuffers& b, uint32_t n, int key_bits, hipStream_t stream);
} // namespace ut
