#include "renderer.h"

#include "settings.h"

#include <rocrand/rocrand.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <atomic>
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

uint64_t pt::host::next_generation() {
  static std::atomic<uint64_t> counter{1};
  return counter.fetch_add(1);
}

const FlatScene &pt_scene::flat() const {
  if (flat_gen_ != generation || flat_mat_gen_ != mat_generation) {
    flat_ = flatten(scene);
    flat_gen_ = generation;
    flat_mat_gen_ = mat_generation;
  }
  return flat_;
}

const Bvh &pt_scene::bvh() const {
  if (bvh_gen_ != generation) {
    bvh_ = build_sah_bvh(flat().tris, tree_settings_);
    bvh_gen_ = generation;
  }
  return bvh_;
}

const PackedScene &pt_scene::shading() const {
  const FlatScene &f = flat();
  if (shade_gen_ != generation) {
    packed_.shade.resize(f.tris.size());
    for (size_t i = 0; i < f.tris.size(); ++i)
      packed_.shade[i] = make_shade_tri(f.tris[i], &f.normals[3 * i], f.tri_mat[i]);
    shade_gen_ = generation;
  }
  if (packed_mat_gen_ != mat_generation || packed_mats_of_gen_ != generation) {
    packed_.mats.clear();
    for (const auto &m : f.materials)
      packed_.mats.push_back(m.to<float>());
    packed_mat_gen_ = mat_generation;
    packed_mats_of_gen_ = generation;
  }
  return packed_;
}

const PackedScene &pt_scene::packed() const {
  const FlatScene &f = flat();
  shading();
  if (packed_gen_ != generation) {
    packed_.bvh = pack_gpu_bvh(bvh(), f.tris);
    packed_gen_ = generation;
  }
  if (packed_mat_gen_ != mat_generation || packed_mats_of_gen_ != generation) {
    packed_.mats.clear();
    for (const auto &m : f.materials)
      packed_.mats.push_back(m.to<float>());
    packed_mat_gen_ = mat_generation;
    packed_mats_of_gen_ = generation;
  }
  return packed_;
}

hipError_t DeviceSceneStore::upload(const pt_scene &scene, hipStream_t s) {
  const FlatScene &f = scene.flat();
  const PackedScene &pk = scene.packed();
  const GpuBvh &g = pk.bvh;
  const auto &mats = pk.mats;
  const auto &st = pk.shade;
  hipError_t e;
  if ((e = nodes.upload(g.nodes.data(), g.nodes.size(), s)) != hipSuccess ||
      (e = tris.upload(g.tris.data(), g.tris.size(), s)) != hipSuccess ||
      (e = shade_tris.upload(st.data(), st.size(), s)) != hipSuccess ||
      (e = materials.upload(mats.data(), mats.size(), s)) != hipSuccess ||
      (e = light_tris.upload(f.light_tris.data(), f.light_tris.size(), s)) !=
          hipSuccess ||
      (e = light_cdf.upload(f.light_cdf.data(), f.light_cdf.size(), s)) !=
          hipSuccess)
    return e;
  dev.nodes = reinterpret_cast<const float4 *>(nodes.ptr);
  dev.tris = reinterpret_cast<const float4 *>(tris.ptr);
  dev.n_nodes = uint32_t(g.nodes.size());
  dev.n_tris = uint32_t(g.tris.size());
  dev.view.tris = shade_tris.ptr;
  dev.view.materials = materials.ptr;
  dev.view.light_tris = light_tris.ptr;
  dev.view.light_cdf = light_cdf.ptr;
  dev.view.n_lights = uint32_t(f.light_tris.size());
  max_depth = g.max_depth;
  return hipStreamSynchronize(s);
}

hipError_t DeviceSceneStore::upload_materials(const pt_scene &scene, hipStream_t s) {
  const FlatScene &f = scene.flat();
  const PackedScene &pk = scene.packed();
  hipError_t e;
  if ((e = materials.upload(pk.mats.data(), pk.mats.size(), s)) != hipSuccess ||
      (e = light_tris.upload(f.light_tris.data(), f.light_tris.size(), s)) !=
          hipSuccess ||
      (e = light_cdf.upload(f.light_cdf.data(), f.light_cdf.size(), s)) != hipSuccess)
    return e;
  dev.view.materials = materials.ptr;
  dev.view.light_tris = light_tris.ptr;
  dev.view.light_cdf = light_cdf.ptr;
  dev.view.n_lights = uint32_t(f.light_tris.size());
  return hipStreamSynchronize(s);
}

pt_renderer::~pt_renderer() {
  if (ev_start)
    (void)hipEventDestroy(ev_start);
  if (ev_stop)
    (void)hipEventDestroy(ev_stop);
  if (ev_fork)
    (void)hipEventDestroy(ev_fork);
  for (auto &pipe : ev_prof)
    for (hipEvent_t e : pipe)
      if (e)
        (void)hipEventDestroy(e);
  if (wf_host_counts)
    (void)hipHostFree(wf_host_counts);
  for (hipStream_t ps : pipe_streams)
    if (ps)
      (void)hipStreamDestroy(ps);
  if (stream)
    (void)hipStreamDestroy(stream);
}

bool pt_renderer::init(int dev, std::string *err) {
  device = dev;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  for (int h = 0; h + 1 < kMaxPipes && e == hipSuccess; ++h)
    e = hipStreamCreateWithFlags(&pipe_streams[h], hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipEventCreate(&ev_start);
  if (e == hipSuccess)
    e = hipEventCreate(&ev_stop);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
  for (int h = 0; h < kMaxPipes && e == hipSuccess; ++h)
    for (int i = 0; i < 3 && e == hipSuccess; ++i)
      e = hipEventCreate(&ev_prof[h][i]);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&n_cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    *err = std::string("HIP init failed: ") + hipGetErrorString(e);
    return false;
  }
  return noalloc_init(dev, err) == PT_OK;
}

int pt_renderer::ensure_scene(const pt_scene &scene, const pt_settings *settings) {
  if (settings)
    scene.use_tree_settings(tree_settings(*settings));
  if (cached_scene == &scene && cached_gen == scene.generation) {
    if (cached_mat_gen == scene.mat_generation)
      return PT_OK;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(scene_store.upload_materials(scene, stream));
    cached_mat_gen = scene.mat_generation;
    return PT_OK;
  }
  const GpuBvh &gb = scene.packed().bvh;
  if (gb.stack_need > uint32_t(gpu_stack_size()))
    return set_error(PT_ERR_UNSUPPORTED,
                     "BVH depth " + std::to_string(gb.max_depth) +
                         " exceeds the traversal stack");
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(scene_store.upload(scene, stream));
  cached_scene = &scene;
  cached_gen = scene.generation;
  cached_mat_gen = scene.mat_generation;
  return PT_OK;
}

int pt_renderer::ensure_sobol(uint32_t dim_bound, uint32_t count) {
  if (sobol_dim_bound == dim_bound && sobol_count == count)
    return PT_OK;
  if (dim_bound == 0 || dim_bound > 20000)
    return set_error(PT_ERR_INVALID, "max_sample_size must be in [1, 20000]");
  const unsigned int *v = nullptr;
  if (rocrand_get_direction_vectors32(&v, ROCRAND_DIRECTION_VECTORS_32_JOEKUO6) !=
          ROCRAND_STATUS_SUCCESS ||
      !v)
    return set_error(PT_ERR_UNSUPPORTED, "Joe-Kuo direction numbers unavailable");
  HIP_TRY(sobol_dirs.upload(v, size_t(dim_bound) * 32, stream));
  HIP_TRY(sobol_vals.ensure(size_t(dim_bound) * count));
  HIP_TRY(gpu::launch_sobol_table(sobol_dirs.ptr, dim_bound, count,
                                  sobol_vals.ptr, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  sobol_dim_bound = dim_bound;
  sobol_count = count;
  return PT_OK;
}

int pt_renderer::sobol_table_to_host(uint32_t dim_bound, uint32_t count,
                                     float *out) {
  HIP_TRY(hipSetDevice(device));
  int rc = ensure_sobol(dim_bound, count);
  if (rc != PT_OK)
    return rc;
  // the device table is dimension-major; the caller gets [sample][dim]
  std::vector<float> cols(size_t(dim_bound) * count);
  HIP_TRY(hipMemcpyAsync(cols.data(), sobol_vals.ptr, cols.size() * 4,
                         hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  for (uint32_t d = 0; d < dim_bound; ++d)
    for (uint32_t s = 0; s < count; ++s)
      out[size_t(s) * dim_bound + d] = cols[size_t(d) * count + s];
  return PT_OK;
}

int pt_renderer::intersect(const pt_scene &scene, const float *rays, uint32_t n,
                           float *t, uint32_t *idx, uint32_t *back) {
  HIP_TRY(hipSetDevice(device));
  int rc = ensure_scene(scene);
  if (rc != PT_OK)
    return rc;
  if (n == 0)
    return PT_OK;
  HIP_TRY(scratch_rays.upload(rays, size_t(n) * 6, stream));
  HIP_TRY(scratch_t.ensure(n));
  HIP_TRY(scratch_idx.ensure(n));
  HIP_TRY(scratch_back.ensure(n));
  HIP_TRY(gpu::launch_intersect(scene_store.dev, scratch_rays.ptr, n,
                                scratch_t.ptr, scratch_idx.ptr,
                                scratch_back.ptr, stream));
  HIP_TRY(hipMemcpyAsync(t, scratch_t.ptr, n * 4, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipMemcpyAsync(idx, scratch_idx.ptr, n * 4, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipMemcpyAsync(back, scratch_back.ptr, n * 4, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return PT_OK;
}

int pt_renderer::render(const pt_sample_spec &spec, const pt_output &out,
                        const pt_scene &scene, uint32_t samples_per,
                        const pt_settings &settings, bool show_progress,
                        bool show_times, double *seconds) {
  (void)show_progress;
  if (samples_per == 0)
    return set_error(PT_ERR_INVALID, "samples_per must be positive");
  uint32_t x_dim, y_dim;
  if (spec.type == PT_SAMPLE_SQUARE_IMAGE) {
    x_dim = spec.x_dim;
    y_dim = spec.y_dim;
  } else if (spec.type == PT_SAMPLE_INITIAL_RAYS ||
             spec.type == PT_SAMPLE_INITIAL_IDX_AND_DIR) {
    if (!spec.rays)
      return set_error(PT_ERR_INVALID, "initial-ray sample spec without rays");
    if (spec.type == PT_SAMPLE_INITIAL_IDX_AND_DIR && !spec.initial_idx)
      return set_error(PT_ERR_INVALID, "InitialIdxAndDir sample spec without triangle indices");
    x_dim = spec.n_rays;
    y_dim = 1;
  } else {
    return set_error(PT_ERR_INVALID, "unknown sample spec type");
  }
  if (x_dim == 0 || y_dim == 0)
    return set_error(PT_ERR_INVALID, "empty image");
  const bool per_step = out.type == PT_OUTPUT_PER_STEP;
  if (per_step) {
    if (!out.per_step || out.n_steps == 0 || out.n_steps > gpu::kMaxOutputSteps)
      return set_error(PT_ERR_INVALID, "per-step output needs 1 to " +
                                           std::to_string(gpu::kMaxOutputSteps) +
                                           " host buffers");
    for (uint32_t i = 0; i < out.n_steps; ++i)
      if (!out.per_step[i])
        return set_error(PT_ERR_INVALID, "null per-step output buffer");
  } else if (!out.data) {
    return set_error(PT_ERR_INVALID, "null output buffer");
  } else if (out.type == PT_OUTPUT_DEVICE_FLOAT_RGB) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, out.data) != hipSuccess ||
        attr.type != hipMemoryTypeDevice || attr.device != device) {
      (void)hipGetLastError();
      return set_error(PT_ERR_INVALID,
                       "DEVICE_FLOAT_RGB output must be device memory of the renderer's GPU");
    }
  } else if (out.type != PT_OUTPUT_BGRA && out.type != PT_OUTPUT_FLOAT_RGB) {
    return set_error(PT_ERR_INVALID, "unknown output type");
  }
  if ((per_step || spec.type != PT_SAMPLE_SQUARE_IMAGE) &&
      settings.kernel_approach != PT_MEGA_KERNEL &&
      !(spec.type == PT_SAMPLE_INITIAL_RAYS && !per_step))
    return set_error(PT_ERR_UNSUPPORTED,
                     "per-step output and initial-intersection sample specs run on the "
                     "MegaKernel integrator");
  if (settings.rng == PT_RNG_SOBEL &&
      (settings.max_sample_size == 0 || settings.max_sample_size > 20000))
    return set_error(PT_ERR_INVALID, "max_sample_size must be in [1, 20000]");
  if (settings.kernel_approach != PT_MEGA_KERNEL &&
      settings.kernel_approach != PT_STREAMING)
    return set_error(PT_ERR_INVALID, "unknown kernel approach");

  HIP_TRY(hipSetDevice(device));
  int rc = ensure_scene(scene, &settings);
  if (rc != PT_OK)
    return rc;
  gpu::RngParams rng{};
  rng.type = settings.rng;
  rng.samples_per = samples_per;
  if (settings.rng == PT_RNG_SOBEL) {
    rc = ensure_sobol(settings.max_sample_size, samples_per);
    if (rc != PT_OK)
      return rc;
    rng.vals = sobol_vals.ptr;
    rng.dim_bound = settings.max_sample_size;
    // reference: src/rng/sobel/sobel.cu:74 (64) and rng_from_sequence_gen.h:101-102
    rng.initial_dim_bound = settings.max_sample_size < 64 ? settings.max_sample_size : 64;
  }
  const uint32_t n_pixels = x_dim * y_dim;
  HIP_TRY(counter.ensure(1));
  HIP_TRY(hipMemsetAsync(counter.ptr, 0, sizeof(unsigned long long), stream));
  if (spec.type != PT_SAMPLE_SQUARE_IMAGE)
    HIP_TRY(scratch_rays.upload(spec.rays, size_t(spec.n_rays) * 6, stream));
  if (spec.type == PT_SAMPLE_INITIAL_IDX_AND_DIR) {
    const uint32_t n_tris = uint32_t(scene.flat().tris.size());
    for (uint32_t i = 0; i < spec.n_rays; ++i)
      if (spec.initial_idx[i] >= n_tris)
        return set_error(PT_ERR_INVALID, "InitialIdxAndDir triangle index " +
                                             std::to_string(spec.initial_idx[i]) +
                                             " out of range");
    HIP_TRY(scratch_idx.upload(spec.initial_idx, spec.n_rays, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));

  stats = pt_render_stats{};
  stats.profiled = profiling ? 1u : 0u;
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipEventRecord(ev_start, stream));
  const float *final_rgb = nullptr;
  if (settings.kernel_approach == PT_MEGA_KERNEL) {
    WorkDivision wd;
    if (!WorkDivision::make(work_division_settings(settings.render_work_division),
                            samples_per, x_dim, y_dim, &wd))
      return set_error(PT_ERR_INVALID, "invalid render work division settings");
    if (wd.block_size > 1024)
      return set_error(PT_ERR_INVALID, "block_size above 1024");
    const size_t image_floats = size_t(wd.num_sample_blocks) * n_pixels * 3;
    HIP_TRY(rgb.ensure(image_floats * (per_step ? out.n_steps : 1)));
    gpu::MegaParams p{};
    p.scene = scene_store.dev;
    p.settings = integrator_settings(settings);
    p.wd = wd;
    p.rng = rng;
    std::memcpy(p.film_to_world, spec.film_to_world, sizeof(p.film_to_world));
    p.initial_rays = spec.type != PT_SAMPLE_SQUARE_IMAGE ? scratch_rays.ptr : nullptr;
    p.initial_idx = spec.type == PT_SAMPLE_INITIAL_IDX_AND_DIR ? scratch_idx.ptr : nullptr;
    p.out_rgb = rgb.ptr;
    p.n_steps = per_step ? out.n_steps : 0;
    p.divide_by_spp = wd.num_sample_blocks == 1;
    p.ray_counter = counter.ptr;
    uint32_t total = wd.total_num_blocks();
    uint32_t chunk = settings.max_blocks_per_launch_gpu ? settings.max_blocks_per_launch_gpu : total;
    for (uint32_t start = 0; start < total; start += chunk) {
      uint32_t end = start + chunk < total ? start + chunk : total;
      if (profiling)
        HIP_TRY(hipEventRecord(ev_prof[0][0], stream));
      HIP_TRY(gpu::launch_megakernel(p, start, end, mega_split, stream));
      if (profiling) {
        HIP_TRY(hipEventRecord(ev_prof[0][1], stream));
        HIP_TRY(hipEventSynchronize(ev_prof[0][1]));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev_prof[0][0], ev_prof[0][1]));
        stats.mega_ms += ms;
      }
      stats.mega_launches++;
    }
    final_rgb = rgb.ptr;
    if (wd.num_sample_blocks != 1) {
      const uint32_t n_images = per_step ? out.n_steps : 1;
      HIP_TRY(rgb2.ensure(size_t(n_pixels) * 3 * n_images));
      for (uint32_t i = 0; i < n_images; ++i)
        HIP_TRY(gpu::launch_reduce_sample_blocks(rgb.ptr + i * image_floats,
                                                 wd.num_sample_blocks, n_pixels, samples_per,
                                                 rgb2.ptr + size_t(i) * n_pixels * 3, stream));
      final_rgb = rgb2.ptr;
    }
  } else {
    uint64_t wf_rays = 0;
    rc = render_streaming(*this, spec, scene_store.dev, samples_per, settings,
                          rng, x_dim, y_dim, &final_rgb, &wf_rays);
    if (rc != PT_OK)
      return rc;
  }
  if (out.type == PT_OUTPUT_BGRA) {
    HIP_TRY(bgra.ensure(n_pixels));
    HIP_TRY(gpu::launch_to_bgra(final_rgb, n_pixels, bgra.ptr, stream));
  }
  HIP_TRY(hipEventRecord(ev_stop, stream));
  // Host outputs go through pinned staging (one DMA copy, then a host
  // memcpy): copying straight into pageable memory took as long as a small
  // render (0.43 ms for a 256 x 256 image).
  const size_t image_bytes = size_t(n_pixels) * (out.type == PT_OUTPUT_BGRA ? 4 : 12);
  const uint32_t n_images = per_step ? out.n_steps : 1;
  const bool host_out = out.type != PT_OUTPUT_DEVICE_FLOAT_RGB;
  if (host_out)
    HIP_TRY(out_stage.ensure(image_bytes * n_images + sizeof(unsigned long long)));
  char *stage = static_cast<char *>(out_stage.ptr);
  if (out.type == PT_OUTPUT_BGRA)
    HIP_TRY(hipMemcpyAsync(stage, bgra.ptr, image_bytes, hipMemcpyDeviceToHost, stream));
  else if (!host_out)
    HIP_TRY(hipMemcpyAsync(out.data, final_rgb, image_bytes, hipMemcpyDeviceToDevice, stream));
  else
    HIP_TRY(hipMemcpyAsync(stage, final_rgb, image_bytes * n_images, hipMemcpyDeviceToHost,
                           stream));
  unsigned long long rays_traced = 0;
  if (host_out)
    HIP_TRY(hipMemcpyAsync(stage + image_bytes * n_images, counter.ptr, sizeof(rays_traced),
                           hipMemcpyDeviceToHost, stream));
  else
    HIP_TRY(hipMemcpyAsync(&rays_traced, counter.ptr, sizeof(rays_traced),
                           hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (host_out) {
    if (per_step)
      for (uint32_t i = 0; i < out.n_steps; ++i)
        std::memcpy(out.per_step[i], stage + image_bytes * i, image_bytes);
    else
      std::memcpy(out.data, stage, image_bytes);
    std::memcpy(&rays_traced, stage + image_bytes * n_images, sizeof(rays_traced));
  }
  auto t1 = std::chrono::steady_clock::now();
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_stop));
  stats.kernel_ms = ms;
  stats.rays = settings.kernel_approach == PT_STREAMING
                   ? stats.ext_rays + stats.shadow_rays
                   : rays_traced;
  stats.samples = uint64_t(samples_per) * n_pixels;
  double secs = std::chrono::duration<double>(t1 - t0).count();
  if (seconds)
    *seconds = secs;
  if (show_times)
    std::printf("render: %.3f ms kernel, %.3f ms total, %llu rays\n", ms,
                secs * 1e3, rays_traced);
  return PT_OK;
}
