// Host loop of the wavefront integrator: shade -> trace (extension, shadow)
// until every sample has finished, then the per-pixel reduction.
#include "renderer.h"
#include "settings.h"

#include "../gpu/trace.h"
#include "../gpu/wavefront.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>

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

int wavefront_alloc(pt_renderer &r, WfPipes &w, const gpu::WfParams &base,
                    uint32_t n_slots, uint64_t total_samples, bool batched,
                    uint32_t trace_blocks) {
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
  HIP_TRY(r.wf_samples.ensure(size_t(total_samples) * 3));
  // r.n_pipes pipelines (an equal share of the slots each, their own stream,
  // queues and counters) while each gets at least one shade block: the
  // memory-bound shade step of one overlaps the ALU-bound traversal of another.
  w.n = int(std::max(1u, std::min({r.n_pipes, uint32_t(kMaxPipes), n_slots / 256})));
  uint32_t first[kMaxPipes + 1] = {0};
  for (int h = 0; h < w.n; ++h) {
    const uint32_t left = n_slots - first[h], pipes_left = uint32_t(w.n - h);
    uint32_t take = (left + pipes_left - 1) / pipes_left;
    take = std::min(left, (take + 255) / 256 * 256);
    first[h + 1] = first[h] + take;
  }
  uint32_t groups[kMaxPipes], caps[kMaxPipes], groups_total = 0;
  size_t queue_total = 0;
  for (int h = 0; h < w.n; ++h) {
    const uint32_t blocks = (first[h + 1] - first[h] + 255) / 256;
    groups[h] = std::min(gpu::kQueueGroups, blocks);
    caps[h] = (blocks + groups[h] - 1) / groups[h] * 256;
    groups_total += groups[h];
    queue_total += size_t(caps[h]) * groups[h];
  }
  HIP_TRY(r.wf_q_ext.ensure(queue_total));
  HIP_TRY(r.wf_q_shadow.ensure(queue_total));
  HIP_TRY(r.wf_counters.ensure(size_t(kMaxPipes) * gpu::kNumCounters));
  HIP_TRY(r.wf_group_counters.ensure(size_t(kMaxPipes) * gpu::kGroupCounterWords));
  const uint32_t trace_threads = trace_blocks * 256;
  HIP_TRY(r.wf_stack_overflow.ensure(size_t(kMaxPipes) * gpu::kOverflowDepth * trace_threads));
  size_t queue_off = 0;
  uint32_t group_base = 0;
  for (int h = 0; h < w.n; ++h) {
    gpu::WfParams &p = w.p[h];
    const uint32_t f = first[h];
    p = base;
    p.total_samples = total_samples;
    p.n_slots = first[h + 1] - f;
    p.st = {r.wf_meta.ptr + f,      r.wf_mult.ptr + f,    r.wf_light.ptr + f,
            r.wf_total.ptr + f,     r.wf_origin.ptr + f,  r.wf_dir.ptr + f,
            r.wf_light_dir.ptr + f, r.wf_hit_ext.ptr + f, r.wf_hit_light.ptr + f,
            batched ? r.wf_scene.ptr + f : nullptr};
    p.q = {r.wf_q_ext.ptr + queue_off, r.wf_q_shadow.ptr + queue_off, caps[h]};
    queue_off += size_t(caps[h]) * groups[h];
    p.counters = r.wf_counters.ptr + size_t(h) * gpu::kNumCounters;
    p.group_counters = r.wf_group_counters.ptr + size_t(h) * gpu::kGroupCounterWords;
    p.n_groups = groups[h];
    p.n_shade_blocks = (p.n_slots + 255) / 256;
    p.group_base = group_base;
    p.n_groups_total = groups_total;
    group_base += groups[h];
    p.out_samples = r.wf_samples.ptr;
    p.stack_overflow = r.wf_stack_overflow.ptr + size_t(h) * gpu::kOverflowDepth * trace_threads;
    p.overflow_stride = trace_threads;
  }
  return PT_OK;
}

// Enqueue one iteration of one pipeline: fold the previous queue counts into
// the totals and zero them, shade, trace the extension and shadow rays (one
// launch), copy the counts to pinned host memory.
static hipError_t enqueue_iteration(pt_renderer &r, const gpu::WfParams &p, int pipe,
                                    bool batched, uint32_t trace_blocks,
                                    bool profile_events) {
  hipStream_t s = r.pipe_stream(pipe);
  hipEvent_t *ev = r.ev_prof[pipe];
  hipError_t e;
  if ((e = gpu::launch_wf_reset(p.counters, p.group_counters, s)) != hipSuccess)
    return e;
  if ((e = gpu::launch_wf_shade(p, batched, s)) != hipSuccess)
    return e;
  if (profile_events && (e = hipEventRecord(ev[0], s)) != hipSuccess)
    return e;
  if ((e = gpu::launch_wf_trace(p, batched, trace_blocks, s)) != hipSuccess)
    return e;
  if (profile_events && (e = hipEventRecord(ev[1], s)) != hipSuccess)
    return e;
  // queue totals of this iteration (written by the extension trace)
  return hipMemcpyAsync(r.wf_host_counts + size_t(pipe) * gpu::kNumCounters, p.counters,
                        sizeof(uint32_t) * gpu::kNumCounters, hipMemcpyDeviceToHost, s);
}

int wavefront_run(pt_renderer &r, WfPipes &w, bool batched, uint32_t trace_blocks,
                  uint32_t iters_per_sync) {
  hipStream_t s0 = r.stream;
  for (int h = 0; h < w.n; ++h) {
    const gpu::WfParams &p = w.p[h];
    HIP_TRY(hipMemsetAsync(p.st.meta, 0, size_t(p.n_slots) * sizeof(uint4), s0));
    HIP_TRY(hipMemsetAsync(p.group_counters, 0, sizeof(uint32_t) * gpu::kGroupCounterWords, s0));
    HIP_TRY(hipMemsetAsync(p.counters, 0, sizeof(uint32_t) * gpu::kNumCounters, s0));
  }
  if (!r.wf_host_counts)
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&r.wf_host_counts),
                          sizeof(uint32_t) * kMaxPipes * gpu::kNumCounters,
                          hipHostMallocDefault));
  // the other pipelines start after everything enqueued so far on stream 0
  // (scene upload, state clears)
  HIP_TRY(hipEventRecord(r.ev_fork, s0));
  for (int h = 1; h < w.n; ++h)
    HIP_TRY(hipStreamWaitEvent(r.pipe_stream(h), r.ev_fork, 0));
  const bool prof = r.profiling;
  if (iters_per_sync == 0)
    iters_per_sync = 1;
  // Profiling needs per-iteration event reads, so it syncs every iteration.
  if (prof)
    iters_per_sync = 1;
  const bool use_graph = r.use_graphs && !prof;
  // Each pipeline runs its own host loop (pipelines 1.. on host threads of
  // their own), so the pipelines drift apart and one's shade step overlaps
  // another's traversal instead of running in lockstep.
  struct PipeResult {
    hipError_t err = hipSuccess;
    std::string what;
    uint64_t launches = 0;
    double trace_ms = 0.0;
  };
  PipeResult res[kMaxPipes];
  auto run_pipe = [&](int h) {
    PipeResult &out = res[h];
    hipStream_t s = r.pipe_stream(h);
    if ((out.err = hipSetDevice(r.device)) != hipSuccess) {
      out.what = "hipSetDevice";
      return;
    }
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      for (uint32_t k = 0; k < iters_per_sync && e == hipSuccess; ++k)
        e = enqueue_iteration(r, w.p[h], h, batched, trace_blocks, false);
      hipError_t e2 = hipStreamEndCapture(s, &graph);
      if (e == hipSuccess && e2 == hipSuccess)
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph)
        (void)hipGraphDestroy(graph);
      if (e != hipSuccess || e2 != hipSuccess) {
        out.err = e != hipSuccess ? e : e2;
        out.what = "graph capture";
        if (exec)
          (void)hipGraphExecDestroy(exec);
        return;
      }
    }
    for (;;) {
      if (use_graph) {
        out.err = hipGraphLaunch(exec, s);
        out.launches += iters_per_sync;
      } else {
        for (uint32_t k = 0; k < iters_per_sync && out.err == hipSuccess; ++k) {
          out.err = enqueue_iteration(r, w.p[h], h, batched, trace_blocks, prof);
          out.launches += 1;
        }
      }
      hipError_t e = hipStreamSynchronize(s);
      if (out.err == hipSuccess)
        out.err = e;
      if (out.err != hipSuccess) {
        out.what = "launch / sync";
        break;
      }
      // counts of the last iteration's queues (pinned host copy in the stream)
      const uint32_t *c = r.wf_host_counts + size_t(h) * gpu::kNumCounters;
      const uint32_t ne = c[gpu::kExtCount], ns = c[gpu::kShadowCount];
      if (prof && ne + ns) {
        float a = 0.f;
        (void)hipEventElapsedTime(&a, r.ev_prof[h][0], r.ev_prof[h][1]);
        out.trace_ms += a;
      }
      // a slot holds a path iff it queued at least one ray
      if (ne + ns == 0)
        break;
    }
    if (exec)
      (void)hipGraphExecDestroy(exec);
  };
  {
    // A pipeline whose thread cannot be started runs on this thread after
    // pipeline 0 (the pipelines are independent, only slower that way).
    std::thread others[kMaxPipes - 1];
    int started = 1;
    try {
      for (; started < w.n; ++started)
        others[started - 1] = std::thread(run_pipe, started);
    } catch (const std::system_error &) {
    }
    run_pipe(0);
    for (int h = started; h < w.n; ++h)
      run_pipe(h);
    for (int h = 1; h < started; ++h)
      others[h - 1].join();
  }
  uint64_t launches = 0;
  for (int h = 0; h < w.n; ++h) {
    if (res[h].err != hipSuccess)
      return set_error(PT_ERR_HIP, "wavefront pipeline " + std::to_string(h) + " " +
                                       res[h].what + ": " + hipGetErrorString(res[h].err));
    launches += res[h].launches;
    r.stats.trace_ms += res[h].trace_ms;
  }
  // every pipeline's last iteration queued nothing, so the totals are complete
  uint32_t c[kMaxPipes * gpu::kNumCounters];
  HIP_TRY(hipMemcpyAsync(c, r.wf_counters.ptr, sizeof(uint32_t) * w.n * gpu::kNumCounters,
                         hipMemcpyDeviceToHost, s0));
  HIP_TRY(hipStreamSynchronize(s0));
  uint64_t ext = 0, shadow = 0;
  uint32_t busy_iters = 0;
  for (int h = 0; h < w.n; ++h) {
    const uint32_t *ch = c + size_t(h) * gpu::kNumCounters;
    ext += uint64_t(ch[gpu::kExtTotal]) | (uint64_t(ch[gpu::kExtTotal + 1]) << 32);
    shadow += uint64_t(ch[gpu::kShadowTotal]) | (uint64_t(ch[gpu::kShadowTotal + 1]) << 32);
    busy_iters += ch[gpu::kBusyIters];
  }
  r.stats.ext_rays = ext;
  r.stats.shadow_rays = shadow;
  // traversal launches over all pipelines (each traces both queues)
  r.stats.busy_trace_launches = busy_iters;
  r.stats.trace_launches = launches;
  r.stats.pipelines = uint32_t(w.n);
  return PT_OK;
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
  const uint32_t trace_blocks = uint32_t(r.n_cus) * r.trace_blocks_for(false);
  gpu::WfParams p{};
  p.scene = scene;
  p.settings = integrator_settings(settings);
  p.rng = rng;
  std::memcpy(p.film_to_world, spec.film_to_world, sizeof(p.film_to_world));
  p.x_dim = x_dim;
  p.y_dim = y_dim;
  WfPipes w;
  int rc = wavefront_alloc(r, w, p, uint32_t(slots64), total, false, trace_blocks);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(r.rgb.ensure(size_t(wd.num_sample_blocks) * n_pixels * 3));
  rc = wavefront_run(r, w, false, trace_blocks, r.iters_per_sync);
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
