// Renderer state behind the C ABI: scene caches, device buffers, launches.
#pragma once

#include "../../../include/path_amd.h"
#include "../gpu/kernels.h"
#include "../gpu/lbvh.h"
#include "../gpu/wavefront.h"
#include "bvh.h"
#include "scene.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <string>

namespace pt {
namespace host {
int set_error(int code, const std::string &msg);
// Device address of the no-allocation entry point's scratch array on
// `device` (current), once per device (gpu/noalloc.hip).
int noalloc_init(int device, std::string *err);
inline int gpu_stack_size() { return gpu::kTraceStackSize; }
// Tree build settings from the accel fields of render settings
// (sbvh/settings.h, bvh/settings.h).
inline BvhSettings tree_settings(const pt_settings &s) {
  BvhSettings b;
  b.traversal_per_intersect_cost = s.traversal_per_intersect_cost;
  b.spatial_splits = s.use_spatial_splits != 0;
  b.overlap_threshold = s.overlap_threshold;
  return b;
}
} // namespace host
} // namespace pt

namespace pt {
namespace host {
// Device-layout arrays of one scene, ready to copy: wave64 BVH, shading
// records, FP32 materials.
struct PackedScene {
  GpuBvh bvh;
  std::vector<ShadeTri<float>> shade;
  std::vector<Material<float>> mats;
};
} // namespace host
} // namespace pt

namespace pt {
namespace host {
// Process-wide unique stamp for scene versions (never reused, so a renderer's
// cache cannot mistake a new scene at a recycled address for an old one).
uint64_t next_generation();
} // namespace host
} // namespace pt

// Caches are rebuilt when `generation` (geometry and materials) or
// `mat_generation` (material values only) changes; distinct scenes may be
// built on different threads at once.
struct pt_scene {
  pt::host::Scene scene;
  // Generation of the scene's derived data (flattened arrays, tree): bumped
  // by geometry edits and by a change of tree build settings.
  mutable uint64_t generation = pt::host::next_generation();
  uint64_t mat_generation = 0;

  // Build settings of the SAH tree; a render with other accel settings
  // rebuilds the tree (as the reference builds its accel per render,
  // include/render/detail/renderer_impl.h:40-73).
  void use_tree_settings(const pt::host::BvhSettings &s) const {
    if (!(s == tree_settings_)) {
      tree_settings_ = s;
      generation = pt::host::next_generation();
    }
  }
  const pt::host::BvhSettings &tree_settings() const { return tree_settings_; }

  const pt::host::FlatScene &flat() const;
  const pt::host::Bvh &bvh() const;
  const pt::host::PackedScene &packed() const;
  // Shading records and materials only (the `bvh` member of the result is
  // filled only once packed() ran): batches whose trees are built on the GPU.
  const pt::host::PackedScene &shading() const;

private:
  mutable uint64_t flat_gen_ = ~0ull, flat_mat_gen_ = ~0ull, bvh_gen_ = ~0ull,
                   shade_gen_ = ~0ull, packed_gen_ = ~0ull, packed_mat_gen_ = ~0ull,
                   packed_mats_of_gen_ = ~0ull;
  mutable pt::host::BvhSettings tree_settings_{pt::host::kGpuSahTraversalCost};
  mutable pt::host::FlatScene flat_;
  mutable pt::host::Bvh bvh_;
  mutable pt::host::PackedScene packed_;
};

namespace pt {
namespace host {

// Grow-only device allocation.
template <class T> struct DeviceBuf {
  T *ptr = nullptr;
  size_t count = 0;
  DeviceBuf() = default;
  DeviceBuf(const DeviceBuf &) = delete;
  DeviceBuf &operator=(const DeviceBuf &) = delete;
  ~DeviceBuf() {
    if (ptr)
      (void)hipFree(ptr);
  }
  hipError_t ensure(size_t n) {
    if (n <= count && ptr)
      return hipSuccess;
    if (ptr)
      (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
    hipError_t e = hipMalloc(&ptr, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess)
      count = n;
    return e;
  }
  hipError_t upload(const T *src, size_t n, hipStream_t s) {
    hipError_t e = ensure(n);
    if (e != hipSuccess || n == 0)
      return e;
    return hipMemcpyAsync(ptr, src, n * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

// Device-resident copy of one scene.
struct DeviceSceneStore {
  DeviceBuf<GpuNode> nodes;
  DeviceBuf<GpuTri> tris;
  DeviceBuf<ShadeTri<float>> shade_tris;
  DeviceBuf<Material<float>> materials;
  DeviceBuf<TriGeom<float>> light_tris;
  DeviceBuf<float> light_cdf;
  gpu::DeviceScene dev;
  uint32_t max_depth = 0;

  hipError_t upload(const pt_scene &scene, hipStream_t s);
  // Material values and the emitter CDF only (geometry unchanged).
  hipError_t upload_materials(const pt_scene &scene, hipStream_t s);
};

// Grow-only pinned host allocation (staging for batched copies).
struct PinnedBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() {
    if (ptr)
      (void)hipHostFree(ptr);
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && ptr)
      return hipSuccess;
    if (ptr)
      (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&ptr, n ? n : 1, hipHostMallocDefault);
    if (e == hipSuccess)
      bytes = n;
    return e;
  }
};

} // namespace host
} // namespace pt

// Independent wavefront pipelines of one render (each its own stream).
constexpr int kMaxPipes = 4;

struct pt_renderer {
  int device = 0;
  hipStream_t stream = nullptr;   // everything, and wavefront pipeline 0
  hipStream_t pipe_streams[kMaxPipes - 1] = {}; // wavefront pipelines 1..
  uint32_t n_pipes = 3; // wavefront pipelines of a render (1..kMaxPipes)
  hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_fork = nullptr;
  pt_render_stats stats{};
  bool profiling = false;
  hipEvent_t ev_prof[kMaxPipes][3] = {};
  hipStream_t pipe_stream(int pipe) const { return pipe == 0 ? stream : pipe_streams[pipe - 1]; }

  // scene cache
  const pt_scene *cached_scene = nullptr;
  uint64_t cached_gen = ~0ull, cached_mat_gen = ~0ull;
  pt::host::DeviceSceneStore scene_store;

  // sampler cache
  uint32_t sobol_dim_bound = 0, sobol_count = 0;
  pt::host::DeviceBuf<uint32_t> sobol_dirs;
  pt::host::DeviceBuf<float> sobol_vals;

  // outputs / scratch
  pt::host::DeviceBuf<float> rgb, rgb2;
  pt::host::DeviceBuf<uint32_t> bgra;
  pt::host::DeviceBuf<unsigned long long> counter;
  pt::host::DeviceBuf<float> scratch_rays;
  pt::host::DeviceBuf<float> scratch_t;
  pt::host::DeviceBuf<uint32_t> scratch_idx, scratch_back;

  // wavefront integrator buffers
  pt::host::DeviceBuf<uint4> wf_meta;
  pt::host::DeviceBuf<uint2> wf_hit_ext, wf_hit_light;
  pt::host::DeviceBuf<float4> wf_mult, wf_light_dir;
  pt::host::DeviceBuf<pt::gpu::F3> wf_light, wf_total, wf_origin, wf_dir;
  pt::host::DeviceBuf<uint32_t> wf_q_ext, wf_q_shadow;
  pt::host::DeviceBuf<int> wf_stack_overflow;
  pt::host::DeviceBuf<uint32_t> wf_counters;
  pt::host::DeviceBuf<uint32_t> wf_group_counters;
  pt::host::DeviceBuf<float> wf_samples;
  pt::host::DeviceBuf<uint32_t> wf_scene;
  uint32_t *wf_host_counts = nullptr; // pinned

  // batched-render buffers (concatenated scenes)
  pt::host::DeviceBuf<pt::host::GpuNode> b_nodes;
  pt::host::DeviceBuf<pt::host::GpuTri> b_tris;
  pt::host::DeviceBuf<pt::ShadeTri<float>> b_shade;
  pt::host::DeviceBuf<pt::Material<float>> b_mats;
  pt::host::DeviceBuf<pt::TriGeom<float>> b_light_tris;
  pt::host::DeviceBuf<float> b_light_cdf;
  pt::host::DeviceBuf<pt::gpu::SceneDesc> b_desc;
  pt::host::DeviceBuf<unsigned long long> b_sample_begin;
  pt::host::DeviceBuf<float> b_gbuffer;
  pt::host::PinnedBuf b_stage_in, b_stage_out;
  pt::host::PinnedBuf out_stage; // host outputs of pt_render
  // the batch whose scenes are on the GPU: (spp, tree builder, then scene
  // generation, material generation and size per item)
  std::vector<uint64_t> batch_key;
  std::vector<pt::gpu::SceneDesc> batch_desc;
  uint64_t batch_pixels = 0, batch_samples = 0;

  // batched ray-sample-spec renders: per-item descriptors and grid prefix
  pt::host::DeviceBuf<pt::gpu::MegaBatchItem> mb_items;
  pt::host::DeviceBuf<uint32_t> mb_blocks;

  // adjoint buffers
  pt::host::DeviceBuf<float> adj_d_image, adj_partials, adj_out;
  pt::host::DeviceBuf<double> adj_group_sums;

  // last Streaming SquareImage render (for the FP64 re-summation)
  uint32_t last_wf_x = 0, last_wf_y = 0, last_wf_spp = 0;
  pt::host::DeviceBuf<double> sum64;

  // radiance-cache render buffers
  pt::host::DeviceBuf<float> nrc_base, nrc_weight, nrc_features, nrc_direct;
  int n_cus = 256;
  // wavefront launch parameters (tunable; results do not depend on them)
  // persistent traversal waves; lanes refill from the wave's queue range.
  // Pipelines x blocks per CU on the 82k-triangle scene (one traversal
  // launch per iteration, both ray queues shared evenly by the waves):
  // 3 x 3 / 4 / 5 / 6 measured 894 / 881 / 867 / 843 Msamples/s, 2 x 6 / 8
  // 870 / 838. Three pipelines of 3 blocks are the 9 traversal blocks (16.5 KB
  // of LDS stack each) that fit a CU's 160 KB at once. Batches of many small
  // scenes prefer more (1024 generated scenes, batch render ms with a pool
  // of 8 M paths: 4 / 6 / 8 blocks 139.6 / 136.1 / 138.0; with 16 M, 6 / 8:
  // 133.7 / 134.3). 0 = those defaults; pt_set_launch_params or the
  // autotuner set a value.
  uint32_t trace_blocks_per_cu = 0;
  uint32_t trace_blocks_for(bool batched) const {
    return trace_blocks_per_cu ? trace_blocks_per_cu : (batched ? 6u : 3u);
  }
  // mega-kernel: several lanes per work-division thread (same image bits)
  bool mega_split = true;
  uint32_t iters_per_sync = 1;
  bool use_graphs = false;
  bool gpu_bvh = false; // batched renders build their trees on the GPU
  int gpu_builder = pt::gpu::kLbvhPloc; // which GPU builder (gpu/lbvh.h LbvhBuilder)
  double last_gpu_bvh_ms = 0.0;     // host wall time of the last GPU tree build
  uint32_t last_gpu_bvh_levels = 0; // levels of its deepest 4-wide tree
  // phases of the last batch that was not resident (host wall ms): scene
  // preparation (flattening, shading records, host trees), packing into the
  // pinned staging buffer, upload, GPU tree builds
  double last_batch_phase_ms[4] = {0.0, 0.0, 0.0, 0.0};
  // GPU tree build scratch
  pt::host::DeviceBuf<uint32_t> adj_overflow;
  pt::host::DeviceBuf<float> lb_scene_box, lb_ibox, lb_lbox;
  pt::host::DeviceBuf<unsigned long long> lb_keys, lb_keys_sorted;
  pt::host::DeviceBuf<uint32_t> lb_vals, lb_vals_sorted, lb_count, lb_visits, lb_counters,
      lb_tri_begin, lb_int_begin;
  pt::host::DeviceBuf<int32_t> lb_left, lb_right, lb_parent_int, lb_parent_leaf, lb_nn;
  pt::host::DeviceBuf<float> lb_clusters[2];
  pt::host::DeviceBuf<pt::gpu::LbvhRoot> lb_frontier[2];
  pt::host::DeviceBuf<char> lb_sort_temp;

  ~pt_renderer();
  bool init(int device, std::string *err);
  // Makes the scene resident; with settings, its tree is first (re)built
  // with their accel fields.
  int ensure_scene(const pt_scene &scene, const pt_settings *settings = nullptr);
  int ensure_sobol(uint32_t dim_bound, uint32_t count);
  int render(const pt_sample_spec &spec, const pt_output &out,
             const pt_scene &scene, uint32_t samples_per,
             const pt_settings &settings, bool show_progress, bool show_times,
             double *seconds);
  int intersect(const pt_scene &scene, const float *rays, uint32_t n, float *t,
                uint32_t *idx, uint32_t *back);
  int sobol_table_to_host(uint32_t dim_bound, uint32_t count, float *out);
};

namespace pt {
namespace host {
// The wavefront integrator's path pool, split into w.n pipelines.
struct WfPipes {
  gpu::WfParams p[kMaxPipes];
  int n = 1;
};
// Allocate the state for n_slots paths and fill the pipelines' parameters
// from `base` (scene, settings, rng, camera, batch, G-buffer).
int wavefront_alloc(pt_renderer &r, WfPipes &w, const gpu::WfParams &base,
                    uint32_t n_slots, uint64_t total_samples, bool batched,
                    uint32_t trace_blocks);
int wavefront_run(pt_renderer &r, WfPipes &w, bool batched, uint32_t trace_blocks,
                  uint32_t iters_per_sync);
int render_batch(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                 uint32_t spp, const pt_settings &settings, double *seconds);
int render_rays_batch(pt_renderer &r, const pt_ray_batch_item *items, uint32_t n,
                      uint32_t spp, uint32_t n_steps, const pt_settings &settings,
                      double *seconds);
int autotune_batch(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                   uint32_t spp, pt_settings &settings, uint32_t out[5]);
int render_adjoint(pt_renderer &r, const pt_scene &scene, const float film_to_world[12],
                   uint32_t x_dim, uint32_t y_dim, uint32_t spp,
                   const pt_settings &settings, const float *d_image, float *grad_albedo,
                   float *grad_emission, double *seconds);
// Wavefront ("Streaming") integrator; writes the final per-pixel FloatRGB
// (already divided by samples_per) into a device buffer owned by the renderer.
int render_streaming(pt_renderer &r, const pt_sample_spec &spec,
                     const gpu::DeviceScene &scene, uint32_t samples_per,
                     const pt_settings &settings, const gpu::RngParams &rng,
                     uint32_t x_dim, uint32_t y_dim, const float **final_rgb,
                     uint64_t *rays_out);
} // namespace host
} // namespace pt
