// Thread -> (pixel, sample range) assignment of the mega-kernel and the
// per-pixel reduction tree.
// reference: src/kernel/work_division.cpp:12-69, include/kernel/work_division_impl.h:10-55,
// include/kernel/work_division_settings.h:8-14
#pragma once

#include "pt_defs.h"

#include <cmath>

namespace pt {

struct WorkDivisionSettings {
  uint32_t block_size = 256;
  uint32_t target_x_block_size = 32;
  uint32_t force_target_samples = 0;
  uint32_t forced_target_samples_per_thread = 8;
  uint32_t base_num_threads = 16384;
  float samples_per_thread_scaling_power = 0.5f;
  uint32_t max_samples_per_thread = 32;
};

PT_INLINE uint32_t closest_power_of_2(uint32_t n) {
  if (n == 0)
    return 1;
  uint32_t lead = 31 - uint32_t(__builtin_clz(n));
  uint32_t smaller = 1u << lead;
  if (lead == 31)
    return smaller;
  uint32_t larger = smaller << 1;
  return (n - smaller <= larger - n) ? smaller : larger;
}

PT_INLINE bool is_power_of_2(uint32_t n) { return n != 0 && (n & (n - 1)) == 0; }

struct ThreadInfo {
  uint32_t x, y, start_sample, end_sample;
  bool exit;
};

struct WorkDivision {
  uint32_t block_size;
  uint32_t base_samples_per_thread;
  uint32_t n_threads_per_unit_extra;
  uint32_t sample_block_size;
  uint32_t x_block_size;
  uint32_t y_block_size;
  uint32_t num_sample_blocks;
  uint32_t num_x_blocks;
  uint32_t num_y_blocks;
  uint32_t x_dim;
  uint32_t y_dim;

  PT_HD uint32_t total_num_blocks() const {
    return num_sample_blocks * num_x_blocks * num_y_blocks;
  }
  PT_HD uint32_t sample_block_idx(uint32_t block_idx) const {
    return block_idx % num_sample_blocks;
  }

  PT_INLINE ThreadInfo get_thread_info(uint32_t block_idx,
                                       uint32_t thread_idx) const {
    uint32_t block_idx_sample = block_idx % num_sample_blocks;
    uint32_t block_idx_location = block_idx / num_sample_blocks;
    uint32_t block_idx_x = block_idx_location % num_x_blocks;
    uint32_t block_idx_y = block_idx_location / num_x_blocks;
    uint32_t thread_idx_sample = thread_idx % sample_block_size;
    uint32_t thread_idx_location = thread_idx / sample_block_size;
    uint32_t thread_idx_x = thread_idx_location % x_block_size;
    uint32_t thread_idx_y = thread_idx_location / x_block_size;
    uint32_t sample_idx = thread_idx_sample + block_idx_sample * sample_block_size;
    ThreadInfo info;
    info.x = thread_idx_x + block_idx_x * x_block_size;
    info.y = thread_idx_y + block_idx_y * y_block_size;
    uint32_t before = sample_idx * base_samples_per_thread;
    uint32_t extra_before = min_(n_threads_per_unit_extra, sample_idx);
    info.start_sample = before + extra_before;
    uint32_t n = base_samples_per_thread +
                 (sample_idx < n_threads_per_unit_extra ? 1u : 0u);
    info.end_sample = info.start_sample + n;
    info.exit = info.x >= x_dim || info.y >= y_dim;
    return info;
  }

  // Returns false (and leaves *out untouched) where the reference's
  // always_assert would abort.
  static bool make(const WorkDivisionSettings &s, uint32_t samples_per,
                   uint32_t x_dim, uint32_t y_dim, WorkDivision *out) {
    if (s.block_size % 64 != 0 || !is_power_of_2(s.block_size) ||
        !is_power_of_2(s.target_x_block_size) || samples_per == 0 ||
        x_dim == 0 || y_dim == 0)
      return false;
    double target;
    if (s.force_target_samples) {
      target = double(s.forced_target_samples_per_thread);
    } else {
      double total = (double(samples_per) * x_dim) * y_dim;
      double ratio = total / double(s.base_num_threads);
      double base = std::pow(ratio, double(s.samples_per_thread_scaling_power));
      target = base < double(s.max_samples_per_thread)
                   ? base
                   : double(s.max_samples_per_thread);
    }
    uint32_t n_threads = closest_power_of_2(
        uint32_t(std::round(double(samples_per) / target)));
    WorkDivision d;
    d.block_size = s.block_size;
    d.base_samples_per_thread = samples_per / n_threads;
    d.n_threads_per_unit_extra = samples_per - d.base_samples_per_thread * n_threads;
    d.sample_block_size = n_threads < s.block_size ? n_threads : s.block_size;
    uint32_t xy = s.block_size / d.sample_block_size;
    d.x_block_size = xy < s.target_x_block_size ? xy : s.target_x_block_size;
    d.y_block_size = xy / d.x_block_size;
    d.num_sample_blocks = n_threads / d.sample_block_size;
    d.num_x_blocks = (x_dim + d.x_block_size - 1) / d.x_block_size;
    d.num_y_blocks = (y_dim + d.y_block_size - 1) / d.y_block_size;
    d.x_dim = x_dim;
    d.y_dim = y_dim;
    if (uint64_t(d.num_sample_blocks) * x_dim * y_dim >= 0xffffffffull)
      return false;
    *out = d;
    return true;
  }
};

// Per-pixel reduction tree of the sub-block partial sums (reference
// include/lib/cuda/reduce.cuh:36-102: shuffle-down halving within each 64-lane
// group, then the same halving over the per-group results). `vals` holds the
// sample_block_size partials of one pixel; the sum is returned. The GPU
// kernels reduce in exactly this order.
template <class V>
PT_INLINE V tree_reduce(V *vals, uint32_t n) {
  uint32_t group = n < 64 ? n : 64;
  for (uint32_t g = 0; g < n; g += group)
    for (uint32_t off = group / 2; off >= 1; off /= 2)
      for (uint32_t i = 0; i < off; ++i)
        vals[g + i] = vals[g + i] + vals[g + i + off];
  uint32_t ng = n / group;
  for (uint32_t off = ng / 2; off >= 1; off /= 2)
    for (uint32_t i = 0; i < off; ++i)
      vals[i * group] = vals[i * group] + vals[(i + off) * group];
  return vals[0];
}

} // namespace pt
