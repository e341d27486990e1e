// Batched multi-scene rendering (training-data generation) and the launch
// autotuner.
#include "renderer.h"
#include "scene_gen.h"
#include "settings.h"

#include "../gpu/trace.h"
#include "../gpu/wavefront.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

using namespace pt;
using namespace pt::host;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error(PT_ERR_HIP, std::string(#expr) + ": " +                 \
                                       hipGetErrorString(e_));                 \
  } while (0)

namespace {

// Flatten + build BVHs of distinct scenes on a host thread pool.
void build_all(const pt_batch_item *items, uint32_t n) {
  std::vector<const pt_scene *> uniq;
  for (uint32_t i = 0; i < n; ++i)
    if (std::find(uniq.begin(), uniq.end(), items[i].scene) == uniq.end())
      uniq.push_back(items[i].scene);
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::atomic<size_t> next{0};
  for (unsigned t = 0; t < hw; ++t)
    pool.emplace_back([&] {
      for (size_t k = next++; k < uniq.size(); k = next++) {
        uniq[k]->flat();
        uniq[k]->bvh();
      }
    });
  for (auto &th : pool)
    th.join();
}

struct Packed {
  std::vector<GpuNode> nodes;
  std::vector<GpuTri> tris;
  std::vector<ShadeTri<float>> shade;
  std::vector<Material<float>> mats;
  std::vector<TriGeom<float>> light_tris;
  std::vector<float> light_cdf;
  std::vector<gpu::SceneDesc> desc;
  std::vector<unsigned long long> sample_begin;
  uint64_t n_pixels = 0;
};

int pack(const pt_batch_item *items, uint32_t n, uint32_t spp, Packed *pk) {
  pk->sample_begin.push_back(0);
  for (uint32_t i = 0; i < n; ++i) {
    const pt_batch_item &it = items[i];
    if (!it.scene || it.x_dim == 0 || it.y_dim == 0 || !it.out_rgb)
      return set_error(PT_ERR_INVALID, "batch item " + std::to_string(i) +
                                           ": null scene/output or empty image");
    const FlatScene &f = it.scene->flat();
    const Bvh &b = it.scene->bvh();
    if (b.max_depth + 1 >= uint32_t(gpu_stack_size()))
      return set_error(PT_ERR_UNSUPPORTED, "batch item " + std::to_string(i) +
                                               ": BVH too deep for the traversal stack");
    GpuBvh g = pack_gpu_bvh(b, f.tris);
    gpu::SceneDesc d{};
    d.node_off = uint32_t(pk->nodes.size());
    d.tri_off = uint32_t(pk->tris.size());
    d.shade_off = uint32_t(pk->shade.size());
    d.mat_off = uint32_t(pk->mats.size());
    d.light_off = uint32_t(pk->light_tris.size());
    d.n_lights = uint32_t(f.light_tris.size());
    d.x_dim = it.x_dim;
    d.y_dim = it.y_dim;
    d.pixel_off = uint32_t(pk->n_pixels);
    d.has_nodes = g.nodes.empty() ? 0u : 1u;
    std::memcpy(d.film_to_world, it.film_to_world, sizeof(d.film_to_world));
    pk->nodes.insert(pk->nodes.end(), g.nodes.begin(), g.nodes.end());
    pk->tris.insert(pk->tris.end(), g.tris.begin(), g.tris.end());
    for (size_t t = 0; t < f.tris.size(); ++t)
      pk->shade.push_back(make_shade_tri(f.tris[t], &f.normals[3 * t], f.tri_mat[t]));
    for (const auto &m : f.materials)
      pk->mats.push_back(m.to<float>());
    pk->light_tris.insert(pk->light_tris.end(), f.light_tris.begin(), f.light_tris.end());
    pk->light_cdf.insert(pk->light_cdf.end(), f.light_cdf.begin(), f.light_cdf.end());
    pk->desc.push_back(d);
    uint64_t px = uint64_t(it.x_dim) * it.y_dim;
    pk->n_pixels += px;
    pk->sample_begin.push_back(pk->sample_begin.back() + px * spp);
  }
  if (pk->n_pixels >= 0xffffffffull)
    return set_error(PT_ERR_INVALID, "batch has too many pixels");
  return PT_OK;
}

} // namespace

namespace pt {
namespace host {

int render_batch(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                 uint32_t spp, const pt_settings &settings, double *seconds) {
  if (n == 0)
    return PT_OK;
  if (spp == 0)
    return set_error(PT_ERR_INVALID, "samples_per must be positive");
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(r.device));
  build_all(items, n);
  Packed pk;
  int rc = pack(items, n, spp, &pk);
  if (rc != PT_OK)
    return rc;
  hipStream_t s = r.stream;
  HIP_TRY(r.b_nodes.upload(pk.nodes.data(), pk.nodes.size(), s));
  HIP_TRY(r.b_tris.upload(pk.tris.data(), pk.tris.size(), s));
  HIP_TRY(r.b_shade.upload(pk.shade.data(), pk.shade.size(), s));
  HIP_TRY(r.b_mats.upload(pk.mats.data(), pk.mats.size(), s));
  HIP_TRY(r.b_light_tris.upload(pk.light_tris.data(), pk.light_tris.size(), s));
  HIP_TRY(r.b_light_cdf.upload(pk.light_cdf.data(), pk.light_cdf.size(), s));
  HIP_TRY(r.b_desc.upload(pk.desc.data(), pk.desc.size(), s));
  HIP_TRY(r.b_sample_begin.upload(pk.sample_begin.data(), pk.sample_begin.size(), s));

  gpu::RngParams rng{};
  rng.type = settings.rng;
  rng.samples_per = spp;
  if (settings.rng == PT_RNG_SOBEL) {
    rc = r.ensure_sobol(settings.max_sample_size, spp);
    if (rc != PT_OK)
      return rc;
    rng.vals = r.sobol_vals.ptr;
    rng.dim_bound = settings.max_sample_size;
    rng.initial_dim_bound = settings.max_sample_size < 64 ? settings.max_sample_size : 64;
  }
  const uint64_t total = pk.sample_begin.back();
  uint64_t slots = settings.max_num_samples_per_launch;
  if (slots == 0)
    return set_error(PT_ERR_INVALID, "max_num_samples_per_launch must be positive");
  slots = std::min<uint64_t>(slots, total);
  const uint32_t trace_blocks = uint32_t(r.n_cus) * r.trace_blocks_per_cu;
  gpu::WfParams p{};
  p.settings = integrator_settings(settings);
  p.rng = rng;
  p.batch.scenes = r.b_desc.ptr;
  p.batch.sample_begin = r.b_sample_begin.ptr;
  p.batch.n_scenes = n;
  p.batch.nodes = reinterpret_cast<const float4 *>(r.b_nodes.ptr);
  p.batch.tris = reinterpret_cast<const float4 *>(r.b_tris.ptr);
  p.batch.shade = r.b_shade.ptr;
  p.batch.materials = r.b_mats.ptr;
  p.batch.light_tris = r.b_light_tris.ptr;
  p.batch.light_cdf = r.b_light_cdf.ptr;
  rc = wavefront_alloc(r, p, uint32_t(slots), total, true, trace_blocks);
  if (rc != PT_OK)
    return rc;
  bool want_gbuffer = false;
  for (uint32_t i = 0; i < n; ++i)
    want_gbuffer |= items[i].out_gbuffer != nullptr;
  if (want_gbuffer) {
    HIP_TRY(r.b_gbuffer.ensure(size_t(pk.n_pixels) * 12));
    HIP_TRY(hipMemsetAsync(r.b_gbuffer.ptr, 0, size_t(pk.n_pixels) * 12 * sizeof(float), s));
    p.gbuffer = r.b_gbuffer.ptr;
  }
  HIP_TRY(r.rgb.ensure(size_t(pk.n_pixels) * 3));
  r.stats = pt_render_stats{};
  r.stats.profiled = r.profiling ? 1u : 0u;
  HIP_TRY(hipEventRecord(r.ev_start, s));
  rc = wavefront_run(r, p, true, trace_blocks, r.iters_per_sync);
  if (rc != PT_OK)
    return rc;
  // per-scene reduction in the single-scene order
  for (uint32_t i = 0; i < n; ++i) {
    const pt_batch_item &it = items[i];
    WorkDivision wd;
    if (!WorkDivision::make(work_division_settings(settings.render_work_division), spp,
                            it.x_dim, it.y_dim, &wd) ||
        wd.block_size > 1024)
      return set_error(PT_ERR_INVALID, "invalid render work division settings");
    size_t off = pk.desc[i].pixel_off;
    size_t px = size_t(it.x_dim) * it.y_dim;
    if (wd.num_sample_blocks == 1) {
      HIP_TRY(gpu::launch_wf_reduce(r.wf_samples.ptr + 3 * off * spp, wd, spp,
                                    r.rgb.ptr + 3 * off, s));
    } else {
      HIP_TRY(r.rgb2.ensure(px * 3 * wd.num_sample_blocks));
      HIP_TRY(gpu::launch_wf_reduce(r.wf_samples.ptr + 3 * off * spp, wd, spp,
                                    r.rgb2.ptr, s));
      HIP_TRY(gpu::launch_reduce_sample_blocks(r.rgb2.ptr, wd.num_sample_blocks,
                                               uint32_t(px), spp, r.rgb.ptr + 3 * off, s));
    }
  }
  HIP_TRY(hipEventRecord(r.ev_stop, s));
  for (uint32_t i = 0; i < n; ++i) {
    size_t off = pk.desc[i].pixel_off;
    size_t px = size_t(items[i].x_dim) * items[i].y_dim;
    HIP_TRY(hipMemcpyAsync(items[i].out_rgb, r.rgb.ptr + 3 * off, px * 12,
                           hipMemcpyDeviceToHost, s));
    if (items[i].out_gbuffer)
      HIP_TRY(hipMemcpyAsync(items[i].out_gbuffer, r.b_gbuffer.ptr + 12 * off, px * 48,
                             hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
  r.stats.kernel_ms = ms;
  r.stats.samples = total;
  r.stats.rays = r.stats.ext_rays + r.stats.shadow_rays;
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}

int autotune_batch(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                   uint32_t spp, pt_settings &settings, uint32_t out[4]) {
  const uint32_t pools[] = {1u << 20, 1u << 21, 1u << 22, 1u << 23};
  const uint32_t blocks[] = {4, 8, 16};
  const uint32_t iters[] = {1, 4};
  double best = 1e30;
  uint32_t bp = settings.max_num_samples_per_launch, bb = r.trace_blocks_per_cu,
           bi = r.iters_per_sync;
  bool saved_graphs = r.use_graphs;
  r.use_graphs = true;
  // warm-up (uploads, Sobol table, code objects)
  int rc = render_batch(r, items, n, spp, settings, nullptr);
  if (rc != PT_OK)
    return rc;
  for (uint32_t pool : pools)
    for (uint32_t b : blocks)
      for (uint32_t it : iters) {
        settings.max_num_samples_per_launch = pool;
        r.trace_blocks_per_cu = b;
        r.iters_per_sync = it;
        rc = render_batch(r, items, n, spp, settings, nullptr);
        if (rc != PT_OK)
          return rc;
        double t = r.stats.kernel_ms;
        if (t < best) {
          best = t;
          bp = pool;
          bb = b;
          bi = it;
        }
      }
  settings.max_num_samples_per_launch = bp;
  r.trace_blocks_per_cu = bb;
  r.iters_per_sync = bi;
  r.use_graphs = true;
  (void)saved_graphs;
  out[0] = bp;
  out[1] = bb;
  out[2] = bi;
  out[3] = 1;
  return PT_OK;
}

} // namespace host
} // namespace pt
