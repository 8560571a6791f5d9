#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut
{
// Entries of the tile-sum scratch exclusive_scan needs for n elements.
uint32_t scan_tmp_size(uint32_t n);
// out[i] = sum(in[0..i-1]); *total (device, optional) = sum(in).  in may equal out.
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, uint32_t* total, hipStream_t s);
} // namespace ut
