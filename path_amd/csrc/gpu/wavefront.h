// Wavefront ("Streaming") integrator: parameter block and launchers.
#pragma once

#include "kernels.h"

namespace pt {
namespace gpu {

// Counter words used by one render (device memory, zeroed by k_wf_reset).
enum WfCounter : uint32_t {
  kExtCount = 0,    // entries in the extension-ray queue
  kShadowCount = 1, // entries in the shadow-ray queue
  kExtHead = 2,     // next extension-ray entry to fetch
  kShadowHead = 3,  // next shadow-ray entry to fetch
  kActive = 4,      // slots holding a path after the last shade pass
  kNumCounters = 8,
};

// Per-slot path state, structure of arrays (one 16-byte vector per slot and
// array so a wave's accesses are contiguous).
struct WfState {
  uint4 *meta;      // location, sample, rng counter, flags
  float4 *mult;     // ray multiplier xyz, light target distance
  float4 *light;    // light-sample multiplier xyz
  float4 *total;    // accumulated radiance xyz
  float4 *origin;   // origin of the pending rays (last vertex / camera)
  float4 *dir;      // direction of the pending continuation ray
  float4 *light_dir; // direction of the pending light ray, target distance
  uint2 *hit_ext;   // packed closest hit of the continuation ray
  uint2 *hit_light; // packed closest hit of the light ray
};

// Ray queues hold slot indices; the rays themselves are read from the state.
struct WfQueues {
  uint32_t *ext;
  uint32_t *shadow;
};

struct WfParams {
  DeviceScene scene;
  IntegratorSettings settings;
  RngParams rng;
  float film_to_world[12];
  uint32_t x_dim, y_dim;
  unsigned long long total_samples;
  uint32_t n_slots;
  WfState st;
  WfQueues q;
  uint32_t *counters;                 // WfCounter words
  unsigned long long *sample_counter; // next sample to start
  unsigned long long *ray_counter;    // rays traced
  float *out_samples;                 // [location * spp + sample] * 3
  int *stack_overflow;                // traversal stack spill area
  uint32_t overflow_stride;           // threads of a traversal launch
};

hipError_t launch_wf_reset(uint32_t *counters, uint32_t mask, hipStream_t s);
hipError_t launch_wf_shade(const WfParams &p, hipStream_t s);
hipError_t launch_wf_trace_ext(const WfParams &p, uint32_t grid_blocks,
                               hipStream_t s);
hipError_t launch_wf_trace_shadow(const WfParams &p, uint32_t grid_blocks,
                                  hipStream_t s);
// Per-pixel sum of the per-sample radiance in the mega-kernel's order
// (sequential per thread range, then the lane tree), divided by spp.
hipError_t launch_wf_reduce(const float *samples, const WorkDivision &wd,
                            uint32_t samples_per, float *out_rgb,
                            hipStream_t s);

} // namespace gpu
} // namespace pt
