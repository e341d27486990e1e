// Wavefront integrator kernels: shade (with in-place path regeneration and
// ray-queue compaction by wave ballot), persistent extension / shadow ray
// traversal over the queues, and the final per-pixel reduction.
//
// Each path slot keeps its state between iterations; a slot whose path ends
// immediately starts the next unstarted sample (one atomic per wave). Rays are
// appended to compact queues, so the traversal kernels run full waves.
// reference semantics: include/render/detail/integrate_image/streaming/run_impl.h:155-460
// (init / intersect / shade loop) with the per-vertex step of
// include/integrate/rendering_equation_iteration.h:9-153.
#include "wavefront.h"

#include "../core/camera.h"
#include "../core/rng.h"
#include "trace.h"
#include "wave.h"

namespace pt {
namespace gpu {

namespace {
constexpr uint32_t kFlagActive = 1u << 24;
constexpr uint32_t kFlagCountEmission = 1u << 25;
constexpr uint32_t kFlagHasNext = 1u << 26;
constexpr uint32_t kFlagLight = 1u << 27;
constexpr uint32_t kItersMask = 0x00ffffffu;

__device__ __forceinline__ uint32_t wave_broadcast_first(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Wave-aggregated append: returns this lane's slot in the queue (valid where
// pred is set). One atomic per wave.
__device__ __forceinline__ uint32_t queue_append(uint32_t *count, bool pred) {
  uint32_t total;
  uint32_t prefix = ballot_prefix(pred, &total);
  uint32_t base = 0;
  if (lane_id() == 0 && total)
    base = atomicAdd(count, total);
  base = wave_broadcast_first(base);
  return base + prefix;
}
} // namespace

__global__ void k_wf_reset(uint32_t *counters, uint32_t mask) {
  uint32_t i = threadIdx.x;
  if (i < kNumCounters && (mask >> i) & 1u)
    counters[i] = 0;
}

hipError_t launch_wf_reset(uint32_t *counters, uint32_t mask, hipStream_t s) {
  hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, s, counters, mask);
  return hipGetLastError();
}

template <class R>
__device__ __forceinline__ R make_rng(const RngParams &rp, uint32_t location,
                                      uint32_t sample, uint32_t counter) {
  if constexpr (sizeof(R) == sizeof(TableRng)) {
    return TableRng{rp.vals, sample * rp.dim_bound, counter, rp.dim_bound};
  } else {
    HashRng h = make_hash_rng(sample + location * rp.samples_per);
    h.counter = counter;
    return h;
  }
}

template <class R> __device__ __forceinline__ uint32_t rng_counter(const R &r) {
  if constexpr (sizeof(R) == sizeof(TableRng))
    return r.dim;
  else
    return r.counter;
}

template <class R>
__device__ __forceinline__ uint32_t rng_start(const RngParams &rp,
                                              uint32_t location) {
  if constexpr (sizeof(R) == sizeof(TableRng))
    return initial_dimension(location, rp.initial_dim_bound);
  else
    return 0;
}

template <class R>
__device__ __forceinline__ void wf_shade(const WfParams &p) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in_range = slot < p.n_slots;
  uint4 meta = in_range ? p.st.meta[slot] : make_uint4(0, 0, 0, 0);
  bool active = in_range && (meta.w & kFlagActive);

  // ---- continue the path in this slot ----
  bool push_ext = false, push_light = false, finished = false;
  V3<float> ext_o = {0.f, 0.f, 0.f}, ext_d = {0.f, 0.f, 0.f}, light_d = {0.f, 0.f, 0.f};
  float light_target = 0.f;
  PathState<float> ns;
  uint32_t location = meta.x, sample = meta.y, counter = meta.z;
  if (active) {
    float4 m = p.st.mult[slot];
    float4 lm = p.st.light[slot];
    float4 tot = p.st.total[slot];
    float4 o = p.st.origin[slot];
    float4 d = p.st.dir[slot];
    PathState<float> st;
    st.iters = meta.w & kItersMask;
    st.count_emission = (meta.w & kFlagCountEmission) ? 1u : 0u;
    st.has_next = (meta.w & kFlagHasNext) ? 1u : 0u;
    st.n_light = (meta.w & kFlagLight) ? 1u : 0u;
    st.ray_has_target = 0;
    st.ray_target = 0.f;
    st.ray_mult = {m.x, m.y, m.z};
    st.light_target = m.w;
    st.light_mult = {lm.x, lm.y, lm.z};
    st.total = {tot.x, tot.y, tot.z};
    Hit<float> light_hit = {0.f, kNoHit, 0}, next_hit = {0.f, kNoHit, 0};
    if (st.n_light)
      light_hit = unpack_hit(p.st.hit_light[slot]);
    if (st.has_next)
      next_hit = unpack_hit(p.st.hit_ext[slot]);
    Ray<float> last = {{o.x, o.y, o.z}, {d.x, d.y, d.z}};
    R rng = make_rng<R>(p.rng, location, sample, counter);
    Ray<float> nl = last, nn = last;
    finished = rendering_equation_iteration(st, rng, last, light_hit, next_hit,
                                            p.settings, p.scene.view, &ns, &nl,
                                            &nn);
    counter = rng_counter(rng);
    if (finished) {
      float *out = p.out_samples +
                   3 * (size_t(location) * p.rng.samples_per + sample);
      out[0] = ns.total.x;
      out[1] = ns.total.y;
      out[2] = ns.total.z;
    } else {
      if (ns.n_light) {
        push_light = true;
        light_d = nl.d;
        light_target = ns.light_target;
        ext_o = nl.o;
      }
      if (ns.has_next) {
        push_ext = true;
        ext_o = nn.o;
        ext_d = nn.d;
      }
    }
  }

  // ---- start a new sample in an empty slot ----
  // Block-aggregated counters: one atomic per block for the sample counter and
  // one per queue (wave ballots + an LDS prefix over the block's 4 waves).
  __shared__ uint32_t s_counts[3][4];
  __shared__ unsigned long long s_sample_base;
  __shared__ uint32_t s_queue_base[2];
  const uint32_t wave = threadIdx.x >> 6;
  bool want_new = in_range && (!active || finished);
  uint32_t n_want;
  uint32_t prefix = ballot_prefix(want_new, &n_want);
  if (lane_id() == 0)
    s_counts[0][wave] = n_want;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = s_counts[0][0] + s_counts[0][1] + s_counts[0][2] + s_counts[0][3];
    s_sample_base = tot ? atomicAdd(p.sample_counter, (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  if (want_new) {
    unsigned long long before = 0;
    for (uint32_t w = 0; w < wave; ++w)
      before += s_counts[0][w];
    unsigned long long s = s_sample_base + before + prefix;
    if (s < p.total_samples) {
      location = uint32_t(s / p.rng.samples_per);
      sample = uint32_t(s % p.rng.samples_per);
      R rng = make_rng<R>(p.rng, location, sample, rng_start<R>(p.rng, location));
      float xo = rng.next();
      float yo = rng.next();
      uint32_t px = location % p.x_dim, py = location / p.x_dim;
      Ray<float> cam = camera_ray(p.film_to_world, float(px) + xo,
                                  float(py) + yo, p.x_dim, p.y_dim);
      counter = rng_counter(rng);
      ns = initial_path_state<float>();
      push_ext = true;
      push_light = false;
      ext_o = cam.o;
      ext_d = cam.d;
    }
  }

  bool holds_path = push_ext || push_light;
  if (holds_path) {
    uint32_t flags = (ns.iters & kItersMask) | kFlagActive |
                     (ns.count_emission ? kFlagCountEmission : 0u) |
                     (ns.has_next ? kFlagHasNext : 0u) |
                     (ns.n_light ? kFlagLight : 0u);
    p.st.meta[slot] = make_uint4(location, sample, counter, flags);
    p.st.mult[slot] = make_float4(ns.ray_mult.x, ns.ray_mult.y, ns.ray_mult.z,
                                  ns.light_target);
    p.st.light[slot] = make_float4(ns.light_mult.x, ns.light_mult.y,
                                   ns.light_mult.z, 0.f);
    p.st.total[slot] = make_float4(ns.total.x, ns.total.y, ns.total.z, 0.f);
    p.st.origin[slot] = make_float4(ext_o.x, ext_o.y, ext_o.z, 0.f);
    p.st.dir[slot] = make_float4(ext_d.x, ext_d.y, ext_d.z, 0.f);
    if (push_light)
      p.st.light_dir[slot] = make_float4(light_d.x, light_d.y, light_d.z, light_target);
  } else if (in_range && (active || want_new)) {
    p.st.meta[slot] = make_uint4(0, 0, 0, 0); // empty slot
  }

  // ---- append rays to the queues (order within the block: wave, lane) ----
  uint32_t n_e, n_s;
  uint32_t pe = ballot_prefix(push_ext, &n_e);
  uint32_t ps = ballot_prefix(push_light, &n_s);
  if (lane_id() == 0) {
    s_counts[1][wave] = n_e;
    s_counts[2][wave] = n_s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t te = s_counts[1][0] + s_counts[1][1] + s_counts[1][2] + s_counts[1][3];
    uint32_t ts = s_counts[2][0] + s_counts[2][1] + s_counts[2][2] + s_counts[2][3];
    s_queue_base[0] = te ? atomicAdd(&p.counters[kExtCount], te) : 0u;
    s_queue_base[1] = ts ? atomicAdd(&p.counters[kShadowCount], ts) : 0u;
  }
  __syncthreads();
  uint32_t be = s_queue_base[0], bs = s_queue_base[1];
  for (uint32_t w = 0; w < wave; ++w) {
    be += s_counts[1][w];
    bs += s_counts[2][w];
  }
  if (push_ext)
    p.q.ext[be + pe] = slot;
  if (push_light)
    p.q.shadow[bs + ps] = slot;
}

// Requires blockDim.x == 256 (block-level counter aggregation).
__global__ __launch_bounds__(256) void k_wf_shade_sobol(WfParams p) {
  wf_shade<TableRng>(p);
}
__global__ __launch_bounds__(256) void k_wf_shade_uniform(WfParams p) {
  wf_shade<HashRng>(p);
}

hipError_t launch_wf_shade(const WfParams &p, hipStream_t s) {
  uint32_t blocks = (p.n_slots + 255) / 256;
  if (p.rng.type == kRngSobel)
    hipLaunchKernelGGL(k_wf_shade_sobol, dim3(blocks), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL(k_wf_shade_uniform, dim3(blocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Persistent traversal: each wave takes 64 queue entries at a time until the
// queue is exhausted (every wave reaches the exit test).
// ---------------------------------------------------------------------------
template <bool kShadow>
__device__ __forceinline__ void wf_trace(const WfParams &p) {
  __shared__ int lds_stack[kLdsStack * 256];
  const uint32_t count = p.counters[kShadow ? kShadowCount : kExtCount];
  uint32_t *head = &p.counters[kShadow ? kShadowHead : kExtHead];
  const uint32_t *q = kShadow ? p.q.shadow : p.q.ext;
  uint2 *hits = kShadow ? p.st.hit_light : p.st.hit_ext;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  // Static schedule: wave w takes 64-entry batches w, w + W, w + 2W, ...
  // (same-address atomics for dynamic fetching cost more than the imbalance).
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  (void)head;
  for (uint32_t base = (gtid >> 6) * 64u; base < count; base += n_waves * 64u) {
    uint32_t i = base + lane_id();
    if (i < count) {
      uint32_t slot = q[i];
      float4 o = p.st.origin[slot];
      float4 d = kShadow ? p.st.light_dir[slot] : p.st.dir[slot];
      Ray<float> ray = {{o.x, o.y, o.z}, {d.x, d.y, d.z}};
      float limit = kShadow ? d.w + 2e-6f : 3.4028235e38f;
      float occl = kShadow ? d.w - 2e-6f : -1.0f;
      LdsStack stack{lds_stack + threadIdx.x, p.stack_overflow + gtid,
                     p.overflow_stride};
      Hit<float> h = trace_closest(p.scene.nodes, p.scene.tris, ray, limit,
                                   occl, stack);
      hits[slot] = pack_hit(h);
    }
  }
}

__global__ __launch_bounds__(256) void k_wf_trace_ext(WfParams p) {
  wf_trace<false>(p);
}
__global__ __launch_bounds__(256) void k_wf_trace_shadow(WfParams p) {
  wf_trace<true>(p);
}

hipError_t launch_wf_trace_ext(const WfParams &p, uint32_t grid_blocks,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_wf_trace_ext, dim3(grid_blocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_wf_trace_shadow(const WfParams &p, uint32_t grid_blocks,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_wf_trace_shadow, dim3(grid_blocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Final reduction with the mega-kernel's grid: thread t of a pixel sums its
// sample range sequentially, then the lanes of the pixel are reduced with the
// same tree as k_megakernel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_wf_reduce(const float *__restrict__ samples,
                                                    WorkDivision wd,
                                                    uint32_t spp,
                                                    float *__restrict__ out) {
  ThreadInfo info = wd.get_thread_info(blockIdx.x, threadIdx.x);
  V3<float> total = {0.f, 0.f, 0.f};
  if (!info.exit) {
    size_t base = size_t(info.x + info.y * wd.x_dim) * spp;
    for (uint32_t s = info.start_sample; s < info.end_sample; ++s) {
      const float *v = samples + 3 * (base + s);
      total += V3<float>{v[0], v[1], v[2]};
    }
  }
  uint32_t sbs = wd.sample_block_size;
  uint32_t group = sbs < 64u ? sbs : 64u;
  total.x = group_reduce_sum(total.x, group);
  total.y = group_reduce_sum(total.y, group);
  total.z = group_reduce_sum(total.z, group);
  bool writer = (threadIdx.x % sbs) == 0 && !info.exit;
  if (sbs > 64u) {
    __shared__ float partial[16][3];
    uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
      partial[w][0] = total.x;
      partial[w][1] = total.y;
      partial[w][2] = total.z;
    }
    __syncthreads();
    if (writer) {
      uint32_t ng = sbs / 64u;
      float v[16][3];
      for (uint32_t g = 0; g < ng; ++g)
        for (int c = 0; c < 3; ++c)
          v[g][c] = partial[w + g][c];
      for (uint32_t off = ng / 2; off >= 1; off /= 2)
        for (uint32_t i = 0; i < off; ++i)
          for (int c = 0; c < 3; ++c)
            v[i][c] = v[i][c] + v[i + off][c];
      total = {v[0][0], v[0][1], v[0][2]};
    }
  }
  if (writer) {
    size_t idx = size_t(wd.sample_block_idx(blockIdx.x)) +
                 size_t(wd.num_sample_blocks) * (size_t(info.x) + size_t(info.y) * wd.x_dim);
    if (wd.num_sample_blocks == 1) {
      float f = float(spp);
      total = {total.x / f, total.y / f, total.z / f};
    }
    out[3 * idx] = total.x;
    out[3 * idx + 1] = total.y;
    out[3 * idx + 2] = total.z;
  }
}

hipError_t launch_wf_reduce(const float *samples, const WorkDivision &wd,
                            uint32_t samples_per, float *out_rgb,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_wf_reduce, dim3(wd.total_num_blocks()), dim3(wd.block_size),
                     0, s, samples, wd, samples_per, out_rgb);
  return hipGetLastError();
}

} // namespace gpu
} // namespace pt
