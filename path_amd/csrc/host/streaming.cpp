// Host loop of the wavefront integrator: shade -> trace (extension, shadow)
// until every sample has finished, then the per-pixel reduction.
#include "renderer.h"
#include "settings.h"

#include "../gpu/trace.h"
#include "../gpu/wavefront.h"

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

int render_streaming(pt_renderer &r, const pt_sample_spec &spec,
                     const gpu::DeviceScene &scene, uint32_t samples_per,
                     const pt_settings &settings, const gpu::RngParams &rng,
                     uint32_t x_dim, uint32_t y_dim, const float **final_rgb,
                     uint64_t *rays_out) {
  if (spec.type != PT_SAMPLE_SQUARE_IMAGE)
    return set_error(PT_ERR_UNSUPPORTED,
                     "the Streaming integrator takes SquareImage sample specs");
  WorkDivision wd;
  if (!WorkDivision::make(work_division_settings(settings.render_work_division),
                          samples_per, x_dim, y_dim, &wd) ||
      wd.block_size > 1024)
    return set_error(PT_ERR_INVALID, "invalid render work division settings");
  const uint64_t n_pixels = uint64_t(x_dim) * y_dim;
  const uint64_t total = n_pixels * samples_per;
  uint64_t slots64 = settings.max_num_samples_per_launch;
  if (slots64 == 0)
    return set_error(PT_ERR_INVALID, "max_num_samples_per_launch must be positive");
  if (slots64 > total)
    slots64 = total;
  const uint32_t n_slots = uint32_t(slots64);
  hipStream_t s = r.stream;
  HIP_TRY(r.wf_meta.ensure(n_slots));
  HIP_TRY(r.wf_hit_ext.ensure(n_slots));
  HIP_TRY(r.wf_hit_light.ensure(n_slots));
  HIP_TRY(r.wf_mult.ensure(n_slots));
  HIP_TRY(r.wf_light.ensure(n_slots));
  HIP_TRY(r.wf_total.ensure(n_slots));
  HIP_TRY(r.wf_origin.ensure(n_slots));
  HIP_TRY(r.wf_dir.ensure(n_slots));
  HIP_TRY(r.wf_light_dir.ensure(n_slots));
  HIP_TRY(r.wf_q_ext.ensure(n_slots));
  HIP_TRY(r.wf_q_shadow.ensure(n_slots));
  HIP_TRY(r.wf_counters.ensure(gpu::kNumCounters));
  HIP_TRY(r.wf_sample_counter.ensure(1));
  HIP_TRY(r.wf_samples.ensure(size_t(total) * 3));
  HIP_TRY(r.rgb.ensure(size_t(wd.num_sample_blocks) * n_pixels * 3));

  gpu::WfParams p{};
  p.scene = scene;
  p.settings = integrator_settings(settings);
  p.rng = rng;
  std::memcpy(p.film_to_world, spec.film_to_world, sizeof(p.film_to_world));
  p.x_dim = x_dim;
  p.y_dim = y_dim;
  p.total_samples = total;
  p.n_slots = n_slots;
  p.st = {r.wf_meta.ptr,   r.wf_mult.ptr, r.wf_light.ptr,
          r.wf_total.ptr,  r.wf_origin.ptr, r.wf_dir.ptr,
          r.wf_light_dir.ptr, r.wf_hit_ext.ptr, r.wf_hit_light.ptr};
  p.q = {r.wf_q_ext.ptr, r.wf_q_shadow.ptr};
  p.counters = r.wf_counters.ptr;
  p.sample_counter = r.wf_sample_counter.ptr;
  p.ray_counter = r.counter.ptr;
  p.out_samples = r.wf_samples.ptr;

  const uint32_t trace_threads = uint32_t(r.n_cus) * 8 * 256;
  HIP_TRY(r.wf_stack_overflow.ensure(size_t(gpu::kOverflowDepth) * trace_threads));
  p.stack_overflow = r.wf_stack_overflow.ptr;
  p.overflow_stride = trace_threads;
  HIP_TRY(hipMemsetAsync(r.wf_meta.ptr, 0, size_t(n_slots) * sizeof(uint4), s));
  HIP_TRY(hipMemsetAsync(r.wf_sample_counter.ptr, 0, sizeof(unsigned long long), s));
  const uint32_t trace_blocks = trace_threads / 256;
  uint64_t launches = 0, rays_total = 0;
  bool pending = false;
  uint32_t words[gpu::kNumCounters] = {};
  for (;;) {
    HIP_TRY(gpu::launch_wf_reset(p.counters,
                                 (1u << gpu::kExtCount) | (1u << gpu::kShadowCount) |
                                     (1u << gpu::kActive),
                                 s));
    HIP_TRY(gpu::launch_wf_shade(p, s));
    HIP_TRY(hipMemcpyAsync(words, p.counters, sizeof(words), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (pending) { // the traversal launches of the previous iteration are done
      float a = 0.f, b = 0.f;
      HIP_TRY(hipEventElapsedTime(&a, r.ev_prof[0], r.ev_prof[1]));
      HIP_TRY(hipEventElapsedTime(&b, r.ev_prof[1], r.ev_prof[2]));
      r.stats.ext_ms += a;
      r.stats.shadow_ms += b;
      pending = false;
    }
    // a slot holds a path iff it queued at least one ray
    if (words[gpu::kExtCount] + words[gpu::kShadowCount] == 0)
      break;
    r.stats.ext_rays += words[gpu::kExtCount];
    r.stats.shadow_rays += words[gpu::kShadowCount];
    rays_total += uint64_t(words[gpu::kExtCount]) + words[gpu::kShadowCount];
    HIP_TRY(gpu::launch_wf_reset(
        p.counters, (1u << gpu::kExtHead) | (1u << gpu::kShadowHead), s));
    if (r.profiling)
      HIP_TRY(hipEventRecord(r.ev_prof[0], s));
    HIP_TRY(gpu::launch_wf_trace_ext(p, trace_blocks, s));
    if (r.profiling)
      HIP_TRY(hipEventRecord(r.ev_prof[1], s));
    HIP_TRY(gpu::launch_wf_trace_shadow(p, trace_blocks, s));
    if (r.profiling) {
      HIP_TRY(hipEventRecord(r.ev_prof[2], s));
      pending = true;
    }
    r.stats.ext_launches++;
    r.stats.shadow_launches++;
    launches += 2;
  }
  HIP_TRY(gpu::launch_wf_reduce(r.wf_samples.ptr, wd, samples_per, r.rgb.ptr, s));
  const float *out = r.rgb.ptr;
  if (wd.num_sample_blocks != 1) {
    HIP_TRY(r.rgb2.ensure(size_t(n_pixels) * 3));
    HIP_TRY(gpu::launch_reduce_sample_blocks(r.rgb.ptr, wd.num_sample_blocks,
                                             uint32_t(n_pixels), samples_per,
                                             r.rgb2.ptr, s));
    out = r.rgb2.ptr;
  }
  r.stats.trace_launches = launches;
  r.stats.rays = rays_total; // overwritten below by the renderer's counter
  *rays_out = rays_total;
  *final_rgb = out;
  return PT_OK;
}

} // namespace host
} // namespace pt
