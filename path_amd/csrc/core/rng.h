// Sample sequences: the Sobol table, the per-pixel dimension walk and a
// counter-based uniform generator.
#pragma once

#include "pt_defs.h"

namespace pt {

// FNV-1a over the 4 little-endian bytes of a 32-bit value.
// reference: include/lib/fnv.h:5-14
PT_INLINE uint32_t fnv_hash(uint32_t in) {
  uint32_t h = 2166136261u;
  for (uint32_t i = 0; i < 4; ++i) {
    h ^= (in >> (8 * i)) & 0xffu;
    h *= 16777619u;
  }
  return h;
}

// 32-bit integer finalizer ("lowbias32", C. Wellons' hash-prospector result,
// constants 0x7feb352d / 0x846ca68b).
PT_INLINE uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Per-dimension digital-shift scramble of the Sobol sequence.
PT_INLINE uint32_t sobol_scramble(uint32_t dim) {
  return mix32(dim ^ 0x5bd1e995u);
}

// Top 24 bits of a 32-bit word as a float in [0, 1 - 2^-24]. Never 1.0, so the
// reference's 1.0 -> 0.0 patch (include/lib/cuda/curand_utils.h:58-66) has no
// counterpart here.
PT_INLINE float u32_to_unit_float(uint32_t x) {
  return float(x >> 8) * 0x1p-24f;
}

// Sobol point `index` (Gray-code order) of one dimension, given that
// dimension's 32 direction numbers (Joe-Kuo "new-joe-kuo-6.21201" set).
PT_INLINE uint32_t sobol_u32(const uint32_t *dirs, uint32_t index) {
  uint32_t gray = index ^ (index >> 1);
  uint32_t x = 0;
  for (uint32_t k = 0; gray != 0; ++k, gray >>= 1)
    if (gray & 1u)
      x ^= dirs[k];
  return x;
}

// Table layout: vals[sample * dimension_bound + dim].
// reference: include/rng/rng_from_sequence_gen.h:26-107 (table walk) and
// src/rng/sobel/sobel.cu:59-70 (dim = i % bound, sample = i / bound).
PT_INLINE float sobol_table_value(const uint32_t *all_dirs, uint32_t dim,
                                  uint32_t sample) {
  return u32_to_unit_float(sobol_u32(all_dirs + 32 * dim, sample) ^
                           sobol_scramble(dim));
}

// Rng over a precomputed table with the per-pixel dimension walk: start
// dimension fnv_hash(location) % initial_dimension_bound, then
// dim = (dim + 1) % dimension_bound after every draw.
// reference: include/rng/rng_from_sequence_gen.h:45-54,88-95
// The GPU table is dimension-major, value(sample, dim) at
// vals[dim * n_samples + sample]: lanes that draw the same dimension for
// consecutive samples (the samples of one pixel) read consecutive words.
struct TableRng {
  const float *vals;
  uint32_t sample;
  uint32_t dim;
  uint32_t dim_bound;
  uint32_t n_samples; // table row length

  PT_INLINE float next() {
    float out = vals[size_t(dim) * n_samples + sample];
    dim = dim + 1 == dim_bound ? 0 : dim + 1;
    return out;
  }
};

PT_INLINE uint32_t initial_dimension(uint32_t location,
                                     uint32_t initial_dim_bound) {
  return fnv_hash(location) % initial_dim_bound;
}

// Counter-based uniform generator standing in for the reference's Uniform rng
// (cuRAND XORWOW on the GPU and std::mt19937 on the CPU, seeded with
// sample_idx + location * samples_per; include/rng/uniform/uniform.h:18-60):
// value = hash(seed, draw count), identical on host and device.
struct HashRng {
  uint32_t key;
  uint32_t counter;

  PT_INLINE float next() {
    uint32_t v = mix32(key ^ mix32(counter * 0x9e3779b9u + 0x632be5abu));
    ++counter;
    return u32_to_unit_float(v);
  }
};

PT_INLINE HashRng make_hash_rng(uint32_t seed) {
  return {mix32(seed ^ 0x68e31da4u), 0u};
}

} // namespace pt
