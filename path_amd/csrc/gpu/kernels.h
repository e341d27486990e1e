// Kernel parameter blocks and host-side launchers of the GPU path.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../core/shading.h"
#include "../core/work_division.h"

namespace pt {
namespace gpu {

// Entries of the per-lane traversal stack (bounds the BVH depth accepted).
constexpr int kTraceStackSize = 64;
// Most per-step images of an OutputPerStep render.
constexpr uint32_t kMaxOutputSteps = 32;

// Device copy of one flattened scene.
struct DeviceScene {
  const float4 *nodes = nullptr; // GpuNode array (4 float4 each)
  const float4 *tris = nullptr;  // GpuTri array in leaf order (3 float4 each)
  SceneView<float> view{};       // shading data in flattened order
  uint32_t n_tris = 0;
  uint32_t n_nodes = 0;
};

struct RngParams {
  uint32_t type; // kRngSobel / kRngUniform
  const float *vals;
  uint32_t dim_bound;
  uint32_t initial_dim_bound;
  uint32_t samples_per;
};

struct MegaParams {
  DeviceScene scene;
  IntegratorSettings settings;
  WorkDivision wd;
  RngParams rng;
  float film_to_world[12];
  const float *initial_rays; // InitialRays / InitialIdxAndDir rays (6 floats) or null
  const uint32_t *initial_idx; // InitialIdxAndDir: triangle hit at distance 1, or null
  float *out_rgb;            // num_sample_blocks * x_dim * y_dim * 3
  uint32_t n_steps;          // OutputPerStep: images (out_rgb holds n_steps of them), else 0
  uint32_t divide_by_spp;    // 1 when num_sample_blocks == 1
  unsigned long long *ray_counter;
};

// One item of a batched ray-sample-spec render (InitialRays /
// InitialIdxAndDir): its scene's offsets into the batch arrays, its work
// division over its rays, and where its rays and outputs start.
struct MegaBatchItem {
  uint32_t node_off, tri_off, shade_off, mat_off, light_off, n_lights;
  uint32_t ray_off;
  uint32_t pad;
  size_t out_off; // floats: n_steps (or 1) images of num_sample_blocks * rays * 3
  WorkDivision wd;
};

struct MegaBatchParams {
  const MegaBatchItem *items;
  const uint32_t *block_begin; // n_items + 1 (first grid block of each item)
  uint32_t n_items;
  uint32_t n_steps;            // per-step images, or 0 for one FloatRGB image
  const float4 *nodes;
  const float4 *tris;
  const ShadeTri<float> *shade;
  const Material<float> *materials;
  const TriGeom<float> *light_tris;
  const float *light_cdf;
  IntegratorSettings settings;
  RngParams rng;
  const float *rays;           // 6 floats per ray, all items
  const uint32_t *initial_idx; // InitialIdxAndDir: triangle index per ray, or null
  float *out;
  unsigned long long *ray_counter;
};

// Path-replay adjoint of a SquareImage render: gradients of sum(dI * I) with
// respect to every material's albedo color (material_albedo) and emission.
constexpr uint32_t kMaxAdjointMaterials = 40; // LDS: 6 * 40 floats per wave
constexpr uint32_t kAdjointBlock = 64;

struct AdjointParams {
  DeviceScene scene;
  IntegratorSettings settings;
  RngParams rng;
  float film_to_world[12];
  uint32_t x_dim, y_dim;
  uint32_t samples_per_thread; // consecutive samples of one pixel per thread
  uint32_t threads_per_pixel;
  uint32_t n_threads;          // pixels * threads_per_pixel
  uint32_t n_mats;
  const float *d_image;        // dLoss/dI, (x + y * x_dim) * 3 + c
  float *partials;             // [block][6 * n_mats]: albedo then emission
  uint32_t single_pass;        // one replay per path (paths of at most
                               // kAdjointMaxVertices vertices)
  uint32_t *overflow;          // set when a single-pass path was longer
};
constexpr uint32_t kAdjointMaxVertices = 16;

// Radiance-cache estimator: the path is traced for one bounce and the
// radiance reflected at the second vertex comes from a learned model. Sample i
// (= pixel * spp + s) records the radiance gathered before that vertex
// (camera-hit emission and the first vertex's light sample), the path
// multiplier reaching the vertex (zero if the bounce escaped or the path was
// terminated) and the vertex features (position, shading normal, albedo).
struct NrcParams {
  DeviceScene scene;
  IntegratorSettings settings;
  RngParams rng;
  float film_to_world[12];
  uint32_t x_dim, y_dim;
  float *base;     // [n][3]
  float *weight;   // [n][3]
  float *features; // [n][9]
};

// Launchers (return hipError_t of the launch).
// out[dim * count + sample] (dimension-major, read by TableRng)
hipError_t launch_sobol_table(const uint32_t *dirs, uint32_t dim_bound,
                              uint32_t count, float *out, hipStream_t s);
hipError_t launch_intersect(const DeviceScene &sc, const float *rays,
                            uint32_t n, float *t, uint32_t *idx, uint32_t *back,
                            hipStream_t s);
// split: run the sample-parallel kernel where it applies (same image bits).
hipError_t launch_megakernel(const MegaParams &p, uint32_t start_block,
                             uint32_t end_block, bool split, hipStream_t s);
// Sum of the num_sample_blocks partials per pixel (in sample-block order),
// divided by samples_per.
hipError_t launch_megakernel_batch(const MegaBatchParams &bp, uint32_t block_size,
                                   uint32_t total_blocks, uint32_t max_blocks_per_launch,
                                   hipStream_t s);
hipError_t launch_reduce_sample_blocks(const float *in, uint32_t nsb,
                                       uint32_t n_pixels, uint32_t samples_per,
                                       float *out, hipStream_t s);
// Per-location reduction of per-sample values over the work division (one
// sample block): out_rgb[3 * (location * num_sample_blocks + block)].x
hipError_t launch_reduce_samples(const WorkDivision &wd, const float *in, uint32_t samples_per,
                                 float *out_rgb, hipStream_t s);
hipError_t launch_to_bgra(const float *rgb, uint32_t n, uint32_t *bgra,
                          hipStream_t s);
hipError_t launch_adjoint(const AdjointParams &p, hipStream_t s);
hipError_t launch_nrc_paths(const NrcParams &p, hipStream_t s);
// Per pixel (pixel-centre camera ray): features of the first hit, the
// emission seen there in base and weight 1 (all zero for a miss).
hipError_t launch_nrc_first_vertex(const NrcParams &p, hipStream_t s);
// out_rgb[pixel] = sum_s (base + weight * max(pred[0..3], 0)) / spp with pred
// rows of 16 floats.
hipError_t launch_nrc_combine(const float *base, const float *weight, const float *pred,
                              uint32_t n_pixels, uint32_t spp, float *out_rgb,
                              hipStream_t s);
// out[i] = scale * sum over blocks of partials[block][i], accumulated in FP64
// in a fixed order: first over contiguous ranges of rows (at most
// kAdjointFinalizeGroups ranges), then over the ranges. group_sums holds
// adjoint_finalize_groups(n_blocks) * n_params doubles; n_params <= 256.
constexpr uint32_t kAdjointFinalizeGroups = 256;
uint32_t adjoint_finalize_groups(uint32_t n_blocks);
hipError_t launch_adjoint_finalize(const float *partials, uint32_t n_blocks,
                                   uint32_t n_params, double scale, double *group_sums,
                                   float *out, hipStream_t s);
// d[i] = 2 (image[i] - target[i]) (device arrays of n floats).
hipError_t launch_l2_gradient(const float *image, const float *target, uint64_t n, float *d,
                              hipStream_t s);

} // namespace gpu
} // namespace pt
