// Host side of the path-replay adjoint (gradients of an image loss with
// respect to material albedo colors and emission).
#include "renderer.h"
#include "settings.h"

#include <chrono>
#include <cstring>

using namespace pt;
using namespace pt::host;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error(PT_ERR_HIP, std::string(#expr) + ": " +                 \
                                       hipGetErrorString(e_));                 \
  } while (0)

namespace pt {
namespace host {

int render_adjoint(pt_renderer &r, const pt_scene &scene, const float film_to_world[12],
                   uint32_t x_dim, uint32_t y_dim, uint32_t spp,
                   const pt_settings &settings, const float *d_image, float *grad_albedo,
                   float *grad_emission, double *seconds) {
  if (x_dim == 0 || y_dim == 0 || spp == 0)
    return set_error(PT_ERR_INVALID, "empty image or zero samples");
  if (!d_image || !grad_albedo || !grad_emission || !film_to_world)
    return set_error(PT_ERR_INVALID, "null argument");
  // The estimator must be polynomial in the parameters for fixed samples:
  // termination may not depend on the path multiplier ...
  if (settings.term_prob != PT_TERM_N_ITERS &&
      settings.term_prob != PT_TERM_CONSTANT)
    return set_error(PT_ERR_UNSUPPORTED,
                     "the adjoint needs a termination probability independent of the path "
                     "multiplier (NIters or Constant)");
  if (settings.rng == PT_RNG_SOBEL &&
      (settings.max_sample_size == 0 || settings.max_sample_size > 20000))
    return set_error(PT_ERR_INVALID, "max_sample_size must be in [1, 20000]");
  // ... and every BSDF a single lobe whose weight is linear in its color.
  const FlatScene &f = scene.flat();
  const uint32_t n_mats = uint32_t(f.materials.size());
  if (n_mats == 0 || n_mats > gpu::kMaxAdjointMaterials)
    return set_error(PT_ERR_UNSUPPORTED, "the adjoint takes 1 to " +
                                             std::to_string(gpu::kMaxAdjointMaterials) +
                                             " materials");
  for (uint32_t i = 0; i < n_mats; ++i)
    if (f.materials[i].type == kDiffuseGlossy || f.materials[i].type == kDiffuseMirror)
      return set_error(PT_ERR_UNSUPPORTED,
                       "material " + std::to_string(i) +
                           ": two-lobe BSDFs pick lobes by albedo and are not differentiable "
                           "with fixed samples");
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(r.device));
  int rc = r.ensure_scene(scene, &settings);
  if (rc != PT_OK)
    return rc;
  gpu::AdjointParams p{};
  p.scene = r.scene_store.dev;
  p.settings = integrator_settings(settings);
  p.rng.type = settings.rng;
  p.rng.samples_per = spp;
  if (settings.rng == PT_RNG_SOBEL) {
    rc = r.ensure_sobol(settings.max_sample_size, spp);
    if (rc != PT_OK)
      return rc;
    p.rng.vals = r.sobol_vals.ptr;
    p.rng.dim_bound = settings.max_sample_size;
    p.rng.initial_dim_bound = settings.max_sample_size < 64 ? settings.max_sample_size : 64;
  }
  std::memcpy(p.film_to_world, film_to_world, sizeof(p.film_to_world));
  p.x_dim = x_dim;
  p.y_dim = y_dim;
  const uint64_t n_pixels = uint64_t(x_dim) * y_dim;
  // about 8192 threads per CU, whole pixels each: 8 rounds of the 4 waves per
  // SIMD that the kernel's VGPRs allow, so a SIMD whose waves end early (short
  // paths) starts new ones instead of idling until the launch's slowest wave
  // ends (256 x 256 at 32 spp, adjoint ms: 1024 threads per CU, one round,
  // 1.14; 2048: 1.00; 4096: 0.94; 8192: 0.92)
  const uint64_t total = n_pixels * spp;
  uint64_t spt = total / (uint64_t(r.n_cus) * 8192);
  spt = spt < 1 ? 1 : (spt > spp ? spp : spt);
  p.samples_per_thread = uint32_t(spt);
  p.threads_per_pixel = uint32_t((spp + spt - 1) / spt);
  const uint64_t n_threads = n_pixels * p.threads_per_pixel;
  if (n_threads >= (1ull << 31))
    return set_error(PT_ERR_INVALID, "image too large for the adjoint launch");
  p.n_threads = uint32_t(n_threads);
  p.n_mats = n_mats;
  const uint32_t n_blocks = uint32_t((n_threads + gpu::kAdjointBlock - 1) / gpu::kAdjointBlock);
  const uint32_t n_params = 6 * n_mats;
  hipStream_t s = r.stream;
  hipPointerAttribute_t attr{};
  const bool d_on_device = hipPointerGetAttributes(&attr, d_image) == hipSuccess &&
                           attr.type == hipMemoryTypeDevice && attr.device == r.device;
  if (!d_on_device) {
    // host d_image: one host memcpy into pinned staging, then one DMA copy
    (void)hipGetLastError();
    const size_t bytes = size_t(n_pixels) * 3 * sizeof(float);
    HIP_TRY(r.out_stage.ensure(bytes));
    std::memcpy(r.out_stage.ptr, d_image, bytes);
    HIP_TRY(r.adj_d_image.ensure(size_t(n_pixels) * 3));
    HIP_TRY(hipMemcpyAsync(r.adj_d_image.ptr, r.out_stage.ptr, bytes, hipMemcpyHostToDevice, s));
  }
  HIP_TRY(r.adj_partials.ensure(size_t(n_blocks) * n_params));
  HIP_TRY(r.adj_out.ensure(n_params));
  HIP_TRY(r.adj_group_sums.ensure(size_t(gpu::adjoint_finalize_groups(n_blocks)) * n_params));
  p.d_image = d_on_device ? d_image : r.adj_d_image.ptr;
  p.partials = r.adj_partials.ptr;
  // one replay per path when the path length is bounded (NIters: at most
  // n_iters + 1 vertices), else radiance replay + back-propagation replay
  p.single_pass = settings.term_prob == PT_TERM_N_ITERS &&
                          settings.n_iters + 1 <= gpu::kAdjointMaxVertices
                      ? 1u
                      : 0u;
  HIP_TRY(r.adj_overflow.ensure(1));
  p.overflow = r.adj_overflow.ptr;
  r.stats = pt_render_stats{};
  std::vector<float> out(n_params);
  for (;;) {
    uint32_t overflow = 0;
    HIP_TRY(hipMemsetAsync(p.overflow, 0, sizeof(uint32_t), s));
    HIP_TRY(hipEventRecord(r.ev_start, s));
    HIP_TRY(gpu::launch_adjoint(p, s));
    HIP_TRY(gpu::launch_adjoint_finalize(r.adj_partials.ptr, n_blocks, n_params,
                                         1.0 / double(spp), r.adj_group_sums.ptr,
                                         r.adj_out.ptr, s));
    HIP_TRY(hipEventRecord(r.ev_stop, s));
    HIP_TRY(hipMemcpyAsync(out.data(), r.adj_out.ptr, n_params * sizeof(float),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&overflow, p.overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!overflow || !p.single_pass)
      break;
    p.single_pass = 0; // a path outran the vertex list: replay twice
  }
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
  r.stats.kernel_ms = ms;
  r.stats.samples = total;
  r.stats.replays = p.single_pass ? 1u : 2u;
  std::memcpy(grad_albedo, out.data(), 3 * n_mats * sizeof(float));
  std::memcpy(grad_emission, out.data() + 3 * n_mats, 3 * n_mats * sizeof(float));
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}

} // namespace host
} // namespace pt
