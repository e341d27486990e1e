// Radiance-cache render: one traced bounce, the learned model for the rest.
#include "mlp.h"
#include "../host/settings.h"

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

extern "C" int pt_render_mlp(pt_renderer *r, pt_mlp *m, const pt_scene *scene,
                             const float film_to_world[12], uint32_t x_dim, uint32_t y_dim,
                             uint32_t samples_per, const pt_settings *settings, float *out_rgb,
                             double *seconds) {
  if (!r || !m || !scene || !film_to_world || !settings || !out_rgb)
    return set_error(PT_ERR_INVALID, "null argument");
  if (x_dim == 0 || y_dim == 0 || samples_per == 0)
    return set_error(PT_ERR_INVALID, "empty image or zero samples");
  if (r->device != m->device)
    return set_error(PT_ERR_INVALID, "renderer and MLP are on different devices");
  const uint64_t n_pixels = uint64_t(x_dim) * y_dim;
  const uint64_t n = n_pixels * samples_per;
  if (n >= (1ull << 32))
    return set_error(PT_ERR_INVALID, "too many samples");
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(r->device));
  int rc = r->ensure_scene(*scene, settings);
  if (rc != PT_OK)
    return rc;
  gpu::NrcParams p{};
  p.scene = r->scene_store.dev;
  p.settings = integrator_settings(*settings);
  p.rng.type = settings->rng;
  p.rng.samples_per = samples_per;
  if (settings->rng == PT_RNG_SOBEL) {
    rc = r->ensure_sobol(settings->max_sample_size, samples_per);
    if (rc != PT_OK)
      return rc;
    p.rng.vals = r->sobol_vals.ptr;
    p.rng.dim_bound = settings->max_sample_size;
    p.rng.initial_dim_bound = settings->max_sample_size < 64 ? settings->max_sample_size : 64;
  }
  std::memcpy(p.film_to_world, film_to_world, sizeof(p.film_to_world));
  p.x_dim = x_dim;
  p.y_dim = y_dim;
  hipStream_t s = r->stream;
  HIP_TRY(r->nrc_base.ensure(n * 3));
  HIP_TRY(r->nrc_weight.ensure(n * 3));
  HIP_TRY(r->nrc_features.ensure(n * 9));
  HIP_TRY(r->rgb.ensure(n_pixels * 3));
  p.base = r->nrc_base.ptr;
  p.weight = r->nrc_weight.ptr;
  p.features = r->nrc_features.ptr;
  HIP_TRY(hipEventRecord(r->ev_start, s));
  HIP_TRY(gpu::launch_nrc_paths(p, s));
  // model queries in chunks of whole pixels (at most 2^22 rows each)
  const uint64_t px_per_chunk = std::max<uint64_t>(1, (1ull << 22) / samples_per);
  for (uint64_t p0 = 0; p0 < n_pixels; p0 += px_per_chunk) {
    const uint64_t p1 = std::min(n_pixels, p0 + px_per_chunk);
    const uint32_t rows = uint32_t((p1 - p0) * samples_per);
    const uint32_t rp = mlp_pad_rows(rows);
    rc = m->ensure_rows(rp, false);
    if (rc != PT_OK)
      return rc;
    HIP_TRY(hipMemsetAsync(m->feat.ptr, 0, size_t(rp) * 9 * sizeof(float), s));
    HIP_TRY(hipMemcpyAsync(m->feat.ptr, r->nrc_features.ptr + p0 * samples_per * 9,
                           size_t(rows) * 9 * sizeof(float), hipMemcpyDeviceToDevice, s));
    rc = m->forward_device(m->feat.ptr, rp, false, s);
    if (rc != PT_OK)
      return rc;
    HIP_TRY(gpu::launch_nrc_combine(r->nrc_base.ptr + p0 * samples_per * 3,
                                    r->nrc_weight.ptr + p0 * samples_per * 3, m->pred.ptr,
                                    uint32_t(p1 - p0), samples_per, r->rgb.ptr + p0 * 3, s));
  }
  HIP_TRY(hipEventRecord(r->ev_stop, s));
  HIP_TRY(hipMemcpyAsync(out_rgb, r->rgb.ptr, n_pixels * 12, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, r->ev_start, r->ev_stop));
  r->stats = pt_render_stats{};
  r->stats.kernel_ms = ms;
  r->stats.samples = n;
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}

extern "C" int pt_render_mlp_first_vertex(pt_renderer *r, pt_mlp *m, const pt_scene *scene,
                                          const float film_to_world[12], uint32_t x_dim,
                                          uint32_t y_dim, uint32_t samples_per,
                                          const pt_settings *settings, float *out_rgb,
                                          double *seconds) {
  if (!r || !m || !scene || !film_to_world || !settings || !out_rgb)
    return set_error(PT_ERR_INVALID, "null argument");
  if (x_dim == 0 || y_dim == 0)
    return set_error(PT_ERR_INVALID, "empty image");
  if (r->device != m->device)
    return set_error(PT_ERR_INVALID, "renderer and MLP are on different devices");
  const uint64_t n_pixels = uint64_t(x_dim) * y_dim;
  if (n_pixels >= (1ull << 32))
    return set_error(PT_ERR_INVALID, "image too large");
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(r->device));
  hipStream_t s = r->stream;
  HIP_TRY(r->nrc_direct.ensure(n_pixels * 3));
  float direct_ms = 0.f;
  if (samples_per) {
    // the tracer's own render, terminated after the first vertex: emission
    // seen by the camera plus the light sample there
    pt_settings d = *settings;
    d.term_prob = PT_TERM_N_ITERS;
    d.n_iters = 0;
    pt_sample_spec spec{};
    spec.type = PT_SAMPLE_SQUARE_IMAGE;
    spec.x_dim = x_dim;
    spec.y_dim = y_dim;
    std::memcpy(spec.film_to_world, film_to_world, sizeof(spec.film_to_world));
    pt_output out{};
    out.type = PT_OUTPUT_DEVICE_FLOAT_RGB;
    out.data = r->nrc_direct.ptr;
    int rc = r->render(spec, out, *scene, samples_per, d, false, false, nullptr);
    if (rc != PT_OK)
      return rc;
    direct_ms = float(r->stats.kernel_ms);
  }
  int rc = r->ensure_scene(*scene, settings);
  if (rc != PT_OK)
    return rc;
  gpu::NrcParams p{};
  p.scene = r->scene_store.dev;
  p.settings = integrator_settings(*settings);
  std::memcpy(p.film_to_world, film_to_world, sizeof(p.film_to_world));
  p.x_dim = x_dim;
  p.y_dim = y_dim;
  HIP_TRY(r->nrc_base.ensure(n_pixels * 3));
  HIP_TRY(r->nrc_weight.ensure(n_pixels * 3));
  HIP_TRY(r->nrc_features.ensure(n_pixels * 9));
  HIP_TRY(r->rgb.ensure(n_pixels * 3));
  p.base = r->nrc_base.ptr;
  p.weight = r->nrc_weight.ptr;
  p.features = r->nrc_features.ptr;
  HIP_TRY(hipEventRecord(r->ev_start, s));
  HIP_TRY(gpu::launch_nrc_first_vertex(p, s));
  const float *base = samples_per ? r->nrc_direct.ptr : r->nrc_base.ptr;
  const uint64_t chunk = 1ull << 22;
  for (uint64_t p0 = 0; p0 < n_pixels; p0 += chunk) {
    const uint64_t p1 = std::min(n_pixels, p0 + chunk);
    const uint32_t rows = uint32_t(p1 - p0);
    const uint32_t rp = mlp_pad_rows(rows);
    rc = m->ensure_rows(rp, false);
    if (rc != PT_OK)
      return rc;
    if (rp > rows)
      HIP_TRY(hipMemsetAsync(m->feat.ptr + size_t(rows) * 9, 0,
                             size_t(rp - rows) * 9 * sizeof(float), s));
    HIP_TRY(hipMemcpyAsync(m->feat.ptr, r->nrc_features.ptr + p0 * 9,
                           size_t(rows) * 9 * sizeof(float), hipMemcpyDeviceToDevice, s));
    rc = m->forward_device(m->feat.ptr, rp, false, s);
    if (rc != PT_OK)
      return rc;
    // one "sample" per pixel: out = base + weight * max(pred, 0)
    HIP_TRY(gpu::launch_nrc_combine(base + p0 * 3, r->nrc_weight.ptr + p0 * 3, m->pred.ptr,
                                    rows, 1, r->rgb.ptr + p0 * 3, s));
  }
  HIP_TRY(hipEventRecord(r->ev_stop, s));
  HIP_TRY(hipMemcpyAsync(out_rgb, r->rgb.ptr, n_pixels * 12, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, r->ev_start, r->ev_stop));
  r->stats = pt_render_stats{};
  r->stats.kernel_ms = ms + direct_ms;
  r->stats.samples = n_pixels * samples_per;
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}
