// Stable LSD radix sort of (64-bit key, 32-bit value) pairs on the device.
//
// 8-bit digits, one pass per digit of bits [0, end_bit). A pass is three
// launches over tiles of 4096 pairs (256 threads x 16): per-tile digit
// histograms (digit-major, so one exclusive scan gives every tile's output
// offset per digit), that scan, and a scatter in which each wave ranks its
// keys among equal digits with eight ballots, so that equal keys keep their
// input order. Per pair and pass the sort reads the key twice and the value
// once and writes both: 32 bytes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pt {
namespace gpu {

// Bytes of temporary storage for n pairs (the per-tile histograms).
size_t radix_sort_temp_bytes(uint32_t n);

// Sorts keys_in/vals_in by bits [0, end_bit) of the keys (1 <= end_bit <=
// 64) into keys_out/vals_out, stably; both pairs of arrays are overwritten
// (the passes alternate between them). temp holds radix_sort_temp_bytes(n).
hipError_t radix_sort_pairs(unsigned long long *keys_in, uint32_t *vals_in,
                            unsigned long long *keys_out, uint32_t *vals_out, uint32_t n,
                            uint32_t end_bit, void *temp, size_t temp_bytes, hipStream_t s);

} // namespace gpu
} // namespace pt
