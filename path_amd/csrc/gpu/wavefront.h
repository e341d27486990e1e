// Wavefront ("Streaming") integrator: parameter block and launchers.
#pragma once

#include "kernels.h"

namespace pt {
namespace gpu {

// Counter words used by one render (device memory, zeroed by k_wf_reset).
enum WfCounter : uint32_t {
  kExtCount = 0,      // entries in the extension-ray queue
  kShadowCount = 1,   // entries in the shadow-ray queue
  kExtTotal = 2,      // 64-bit running sum of kExtCount (words 2, 3)
  kShadowTotal = 4,   // 64-bit running sum of kShadowCount (words 4, 5)
  kBusyIters = 6,     // iterations that queued at least one ray
  kNumCounters = 8,
};

// Queue groups. Shade block b belongs to group b % n_groups (n_groups =
// min(kQueueGroups, shade blocks)): it appends its rays to the group's queue
// segment and starts new samples from the group's share of the sample range
// (shares proportional to the group's blocks). Each group's counters sit in
// their own 4 KB: device-scope atomics on one address serialise (about 12 ns
// each measured on MI355X, i.e. ~200 us per shade step for one counter hit by
// all 16k blocks), so one counter per group spreads them over 16 addresses.
constexpr uint32_t kQueueGroups = 16;
constexpr uint32_t kGroupStride = 1024; // uint32 words between group counters
enum GroupCounter : uint32_t {
  kGroupExt = 0,    // entries in the group's extension-ray segment
  kGroupShadow = 1, // entries in the group's shadow-ray segment
  kGroupSample = 2, // 64-bit: samples the group has started
  kGroupKinds = 3,
};
constexpr size_t kGroupCounterWords = size_t(kGroupKinds) * kQueueGroups * kGroupStride;

__host__ __device__ inline uint32_t *group_counter(uint32_t *base, uint32_t kind,
                                                   uint32_t g) {
  return base + (size_t(kind) * kQueueGroups + g) * kGroupStride;
}

// Three packed floats (12 bytes: one dwordx3 access per lane).
struct F3 {
  float x, y, z;
};

// Per-slot path state, structure of arrays (one 16- or 12-byte record per
// slot and array, so a wave's accesses are contiguous).
struct WfState {
  uint4 *meta;       // location, sample, rng counter, flags
  float4 *mult;      // ray multiplier xyz, light target distance
  F3 *light;         // light-sample multiplier
  F3 *total;         // accumulated radiance
  F3 *origin;        // origin of the pending rays (last vertex / camera)
  F3 *dir;           // direction of the pending continuation ray
  float4 *light_dir; // direction of the pending light ray, target distance
  uint2 *hit_ext;    // packed closest hit of the continuation ray
  uint2 *hit_light;  // packed closest hit of the light ray
  uint32_t *scene;   // scene of the slot (batched renders only)
};

// Ray queues hold slot indices; the rays themselves are read from the state.
// Group g's entries start at g * group_cap; the traversal kernels read the
// group counts and concatenate the segments.
struct WfQueues {
  uint32_t *ext;
  uint32_t *shadow;
  uint32_t group_cap;
};

// One scene of a batched render. Offsets index the concatenated arrays of all
// scenes (nodes in GpuNode units, triangles in GpuTri units, etc.).
struct SceneDesc {
  uint32_t node_off, tri_off, shade_off, mat_off;
  uint32_t light_off, n_lights, x_dim, y_dim;
  uint32_t pixel_off, has_nodes, pad0, pad1;
  float film_to_world[12];
};

struct WfBatch {
  const SceneDesc *scenes;
  const unsigned long long *sample_begin; // n_scenes + 1 prefix of pixels*spp
  uint32_t n_scenes;
  const float4 *nodes, *tris;             // concatenated BVHs
  const ShadeTri<float> *shade;
  const Material<float> *materials;
  const TriGeom<float> *light_tris;
  const float *light_cdf;
};

struct WfParams {
  DeviceScene scene; // single-scene renders
  WfBatch batch;     // batched renders
  IntegratorSettings settings;
  RngParams rng;
  float film_to_world[12];
  uint32_t x_dim, y_dim;
  unsigned long long total_samples;
  uint32_t n_slots;
  WfState st;
  WfQueues q;
  uint32_t *counters;       // WfCounter words
  uint32_t *group_counters; // GroupCounter words per group (group_counter())
  uint32_t n_groups;        // min(kQueueGroups, shade blocks)
  uint32_t n_shade_blocks;
  // Samples are dealt over the groups of every pipeline of a render (the
  // path pool can be split into independent pipelines on separate streams):
  // this launch's group g is group group_base + g of n_groups_total.
  uint32_t group_base, n_groups_total;
  float *out_samples;                 // [pixel * spp + sample] * 3
  float *gbuffer; // first hit of sample 0 per pixel: position, normal,
                  // albedo, emission (12 floats) or null
  int *stack_overflow;                // traversal stack spill area
  uint32_t overflow_stride;           // threads of a traversal launch
};

// Fold the queue counts of the previous iteration into the running totals and
// zero them and the group queue counters (one block; safe inside a captured
// graph).
hipError_t launch_wf_reset(uint32_t *counters, uint32_t *group_counters, hipStream_t s);
hipError_t launch_wf_shade(const WfParams &p, bool batched, hipStream_t s);
// Closest hits of the iteration's extension rays and occlusion of its shadow
// rays, both queues in one launch; also writes the iteration's queue totals
// into counters[kExtCount], counters[kShadowCount].
hipError_t launch_wf_trace(const WfParams &p, bool batched, uint32_t grid_blocks,
                           hipStream_t s);
// Per-pixel sum of the per-sample radiance in the mega-kernel's order
// (sequential per thread range, then the lane tree), divided by spp.
hipError_t launch_wf_reduce(const float *samples, const WorkDivision &wd,
                            uint32_t samples_per, float *out_rgb,
                            hipStream_t s);
// out[pixel] = (sum over the pixel's samples, sequential in FP64) / spp
hipError_t launch_sum_samples_f64(const float *samples, uint32_t n_pixels,
                                  uint32_t samples_per, double *out, hipStream_t s);

} // namespace gpu
} // namespace pt
