// Host loop of the wavefront integrator: shade -> trace (extension, shadow)
// until every sample has finished, then the per-pixel reduction.
#include "renderer.h"
#include "settings.h"

#include "../gpu/trace.h"
#include "../gpu/wavefront.h"

#include <algorithm>
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

int wavefront_alloc(pt_renderer &r, gpu::WfParams &p, uint32_t n_slots,
                    uint64_t total_samples, bool batched, uint32_t trace_blocks) {
  HIP_TRY(r.wf_meta.ensure(n_slots));
  HIP_TRY(r.wf_hit_ext.ensure(n_slots));
  HIP_TRY(r.wf_hit_light.ensure(n_slots));
  HIP_TRY(r.wf_mult.ensure(n_slots));
  HIP_TRY(r.wf_light.ensure(n_slots));
  HIP_TRY(r.wf_total.ensure(n_slots));
  HIP_TRY(r.wf_origin.ensure(n_slots));
  HIP_TRY(r.wf_dir.ensure(n_slots));
  HIP_TRY(r.wf_light_dir.ensure(n_slots));
  if (batched)
    HIP_TRY(r.wf_scene.ensure(n_slots));
  // queue segments: one per group, sized for the group's shade blocks
  const uint32_t shade_blocks = (n_slots + 255) / 256;
  const uint32_t n_groups = std::min(gpu::kQueueGroups, shade_blocks);
  const uint32_t group_cap = (shade_blocks + n_groups - 1) / n_groups * 256;
  HIP_TRY(r.wf_q_ext.ensure(size_t(group_cap) * n_groups));
  HIP_TRY(r.wf_q_shadow.ensure(size_t(group_cap) * n_groups));
  HIP_TRY(r.wf_counters.ensure(gpu::kNumCounters));
  HIP_TRY(r.wf_group_counters.ensure(gpu::kGroupCounterWords));
  HIP_TRY(r.wf_samples.ensure(size_t(total_samples) * 3));
  const uint32_t trace_threads = trace_blocks * 256;
  HIP_TRY(r.wf_stack_overflow.ensure(size_t(gpu::kOverflowDepth) * trace_threads));
  p.total_samples = total_samples;
  p.n_slots = n_slots;
  p.st = {r.wf_meta.ptr,     r.wf_mult.ptr,   r.wf_light.ptr,
          r.wf_total.ptr,    r.wf_origin.ptr, r.wf_dir.ptr,
          r.wf_light_dir.ptr, r.wf_hit_ext.ptr, r.wf_hit_light.ptr,
          batched ? r.wf_scene.ptr : nullptr};
  p.q = {r.wf_q_ext.ptr, r.wf_q_shadow.ptr, group_cap};
  p.counters = r.wf_counters.ptr;
  p.group_counters = r.wf_group_counters.ptr;
  p.n_groups = n_groups;
  p.n_shade_blocks = shade_blocks;
  p.out_samples = r.wf_samples.ptr;
  p.stack_overflow = r.wf_stack_overflow.ptr;
  p.overflow_stride = trace_threads;
  return PT_OK;
}

// Enqueue one iteration: fold the previous queue counts into the totals and
// zero them, shade, trace extension and shadow rays, copy the counts to
// pinned host memory.
static hipError_t enqueue_iteration(pt_renderer &r, const gpu::WfParams &p,
                                    bool batched, uint32_t trace_blocks,
                                    bool profile_events) {
  hipStream_t s = r.stream;
  hipError_t e;
  if ((e = gpu::launch_wf_reset(p.counters, p.group_counters, s)) != hipSuccess)
    return e;
  if ((e = gpu::launch_wf_shade(p, batched, s)) != hipSuccess)
    return e;
  if (profile_events && (e = hipEventRecord(r.ev_prof[0], s)) != hipSuccess)
    return e;
  if ((e = gpu::launch_wf_trace_ext(p, batched, trace_blocks, s)) != hipSuccess)
    return e;
  if (profile_events && (e = hipEventRecord(r.ev_prof[1], s)) != hipSuccess)
    return e;
  if ((e = gpu::launch_wf_trace_shadow(p, batched, trace_blocks, s)) != hipSuccess)
    return e;
  if (profile_events && (e = hipEventRecord(r.ev_prof[2], s)) != hipSuccess)
    return e;
  // queue totals of this iteration (written by the extension trace)
  return hipMemcpyAsync(r.wf_host_counts, p.counters, sizeof(uint32_t) * gpu::kNumCounters,
                        hipMemcpyDeviceToHost, s);
}

int wavefront_run(pt_renderer &r, gpu::WfParams &p, bool batched,
                  uint32_t trace_blocks, uint32_t iters_per_sync) {
  hipStream_t s = r.stream;
  HIP_TRY(hipMemsetAsync(p.st.meta, 0, size_t(p.n_slots) * sizeof(uint4), s));
  HIP_TRY(hipMemsetAsync(p.group_counters, 0, sizeof(uint32_t) * gpu::kGroupCounterWords, s));
  HIP_TRY(hipMemsetAsync(p.counters, 0, sizeof(uint32_t) * gpu::kNumCounters, s));
  if (!r.wf_host_counts)
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&r.wf_host_counts),
                          sizeof(uint32_t) * gpu::kNumCounters, hipHostMallocDefault));
  const bool prof = r.profiling;
  if (iters_per_sync == 0)
    iters_per_sync = 1;
  // Profiling needs per-iteration event reads, so it syncs every iteration.
  if (prof)
    iters_per_sync = 1;
  hipGraphExec_t exec = nullptr;
  bool use_graph = r.use_graphs && !prof;
  if (use_graph) {
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (uint32_t k = 0; k < iters_per_sync && e == hipSuccess; ++k)
      e = enqueue_iteration(r, p, batched, trace_blocks, false);
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph)
        (void)hipGraphDestroy(graph);
      return set_error(PT_ERR_HIP, std::string("graph capture failed: ") +
                                       hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
      return set_error(PT_ERR_HIP, std::string("hipGraphInstantiate: ") +
                                       hipGetErrorString(e));
  }
  uint64_t launches = 0;
  int rc = PT_OK;
  for (;;) {
    if (use_graph) {
      hipError_t e = hipGraphLaunch(exec, s);
      if (e != hipSuccess) {
        rc = set_error(PT_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        break;
      }
      launches += 2 * iters_per_sync;
    } else {
      for (uint32_t k = 0; k < iters_per_sync; ++k) {
        hipError_t e = enqueue_iteration(r, p, batched, trace_blocks, prof);
        if (e != hipSuccess) {
          rc = set_error(PT_ERR_HIP, hipGetErrorString(e));
          break;
        }
        launches += 2;
      }
      if (rc != PT_OK)
        break;
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      rc = set_error(PT_ERR_HIP, std::string("sync: ") + hipGetErrorString(e));
      break;
    }
    // counts of the last iteration's queues (pinned host copy in the stream)
    uint32_t ne = r.wf_host_counts[gpu::kExtCount];
    uint32_t ns = r.wf_host_counts[gpu::kShadowCount];
    if (prof && ne + ns) {
      float a = 0.f, b = 0.f;
      (void)hipEventElapsedTime(&a, r.ev_prof[0], r.ev_prof[1]);
      (void)hipEventElapsedTime(&b, r.ev_prof[1], r.ev_prof[2]);
      r.stats.ext_ms += a;
      r.stats.shadow_ms += b;
    }
    // a slot holds a path iff it queued at least one ray
    if (ne + ns == 0)
      break;
  }
  if (exec)
    (void)hipGraphExecDestroy(exec);
  if (rc != PT_OK)
    return rc;
  // the last iteration queued nothing, so the totals are complete
  uint32_t c[gpu::kNumCounters];
  HIP_TRY(hipMemcpyAsync(c, p.counters, sizeof(c), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  uint64_t ext = uint64_t(c[gpu::kExtTotal]) | (uint64_t(c[gpu::kExtTotal + 1]) << 32);
  uint64_t shadow =
      uint64_t(c[gpu::kShadowTotal]) | (uint64_t(c[gpu::kShadowTotal + 1]) << 32);
  r.stats.ext_rays = ext;
  r.stats.shadow_rays = shadow;
  r.stats.ext_launches = c[gpu::kBusyIters];
  r.stats.shadow_launches = c[gpu::kBusyIters];
  r.stats.trace_launches = launches;
  return rc;
}

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
  const uint32_t trace_blocks = uint32_t(r.n_cus) * r.trace_blocks_per_cu;
  gpu::WfParams p{};
  p.scene = scene;
  p.settings = integrator_settings(settings);
  p.rng = rng;
  std::memcpy(p.film_to_world, spec.film_to_world, sizeof(p.film_to_world));
  p.x_dim = x_dim;
  p.y_dim = y_dim;
  int rc = wavefront_alloc(r, p, uint32_t(slots64), total, false, trace_blocks);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(r.rgb.ensure(size_t(wd.num_sample_blocks) * n_pixels * 3));
  rc = wavefront_run(r, p, false, trace_blocks, r.iters_per_sync);
  if (rc != PT_OK)
    return rc;
  hipStream_t s = r.stream;
  HIP_TRY(gpu::launch_wf_reduce(r.wf_samples.ptr, wd, samples_per, r.rgb.ptr, s));
  const float *out = r.rgb.ptr;
  if (wd.num_sample_blocks != 1) {
    HIP_TRY(r.rgb2.ensure(size_t(n_pixels) * 3));
    HIP_TRY(gpu::launch_reduce_sample_blocks(r.rgb.ptr, wd.num_sample_blocks,
                                             uint32_t(n_pixels), samples_per,
                                             r.rgb2.ptr, s));
    out = r.rgb2.ptr;
  }
  *final_rgb = out;
  *rays_out = r.stats.ext_rays + r.stats.shadow_rays;
  r.last_wf_x = x_dim;
  r.last_wf_y = y_dim;
  r.last_wf_spp = samples_per;
  return PT_OK;
}

} // namespace host
} // namespace pt
