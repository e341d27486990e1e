// Wavefront integrator kernels: shade (with in-place path regeneration and
// ray-queue compaction by wave ballot), persistent extension / shadow ray
// traversal over the queues, and the final per-pixel reduction.
//
// Each path slot keeps its state between iterations; a slot whose path ends
// immediately starts the next unstarted sample. Rays are appended to compact
// queues of slot indices, so the traversal kernels run full waves. Counters
// are aggregated per block (one atomic per block and counter): same-address
// device-scope atomics are the expensive operation on this chip.
//
// Batched renders (many scenes in one launch sequence) are the same kernels
// instantiated with kBatch: every slot also carries its scene, whose BVH,
// shading records, materials and lights are addressed through offsets into
// concatenated arrays.
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
} // namespace

__global__ void k_wf_reset(uint32_t *counters, uint32_t *group_counters) {
  const uint32_t t = threadIdx.x;
  if (t < 2 * kQueueGroups)
    *group_counter(group_counters, t < kQueueGroups ? kGroupExt : kGroupShadow,
                   t % kQueueGroups) = 0;
  if (t != 0)
    return;
  const uint32_t ne = counters[kExtCount], ns = counters[kShadowCount];
  auto *tot = reinterpret_cast<unsigned long long *>(counters);
  tot[kExtTotal / 2] += ne;
  tot[kShadowTotal / 2] += ns;
  counters[kBusyIters] += (ne + ns) != 0u;
  counters[kExtCount] = 0;
  counters[kShadowCount] = 0;
}

hipError_t launch_wf_reset(uint32_t *counters, uint32_t *group_counters, hipStream_t s) {
  hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, s, counters, group_counters);
  return hipGetLastError();
}

// Samples are dealt to the groups in chunks of kSampleChunk consecutive
// samples, round robin, so every group sweeps the whole image at the same
// pace (contiguous shares measured 8% slower on the 82k-triangle scene: the
// groups whose image band has the longer paths set the tail).
constexpr uint32_t kSampleChunk = 256;

// Number of samples group g (of this pipeline) starts.
__device__ __forceinline__ unsigned long long group_sample_count(const WfParams &p,
                                                                 uint32_t g) {
  const unsigned long long round = (unsigned long long)p.n_groups_total * kSampleChunk;
  const unsigned long long rem = p.total_samples % round;
  const unsigned long long first = (unsigned long long)(p.group_base + g) * kSampleChunk;
  const unsigned long long tail =
      rem > first ? (rem - first < kSampleChunk ? rem - first : kSampleChunk) : 0ull;
  return p.total_samples / round * kSampleChunk + tail;
}

// Global index of the i-th sample of group g (of this pipeline).
__device__ __forceinline__ unsigned long long group_sample(const WfParams &p, uint32_t g,
                                                           unsigned long long i) {
  return i / kSampleChunk * ((unsigned long long)p.n_groups_total * kSampleChunk) +
         (unsigned long long)(p.group_base + g) * kSampleChunk + i % kSampleChunk;
}

template <class R>
__device__ __forceinline__ R make_rng(const RngParams &rp, uint32_t location,
                                      uint32_t sample, uint32_t counter) {
  if constexpr (std::is_same_v<R, TableRng>) {
    return TableRng{rp.vals, sample, counter, rp.dim_bound, rp.samples_per};
  } else {
    HashRng h = make_hash_rng(sample + location * rp.samples_per);
    h.counter = counter;
    return h;
  }
}

template <class R> __device__ __forceinline__ uint32_t rng_counter(const R &r) {
  if constexpr (std::is_same_v<R, TableRng>)
    return r.dim;
  else
    return r.counter;
}

template <class R>
__device__ __forceinline__ uint32_t rng_start(const RngParams &rp,
                                              uint32_t location) {
  if constexpr (std::is_same_v<R, TableRng>)
    return initial_dimension(location, rp.initial_dim_bound);
  else
    return 0;
}

// Scene-dependent addressing of one slot.
struct SlotScene {
  SceneView<float> view;
  const float *film_to_world;
  uint32_t x_dim, y_dim, pixel_off;
};

template <bool kBatch>
__device__ __forceinline__ SlotScene slot_scene(const WfParams &p,
                                                uint32_t scene) {
  SlotScene s;
  if constexpr (kBatch) {
    const SceneDesc &d = p.batch.scenes[scene];
    s.view.tris = p.batch.shade + d.shade_off;
    s.view.materials = p.batch.materials + d.mat_off;
    s.view.light_tris = p.batch.light_tris + d.light_off;
    s.view.light_cdf = p.batch.light_cdf + d.light_off;
    s.view.n_lights = d.n_lights;
    s.film_to_world = d.film_to_world;
    s.x_dim = d.x_dim;
    s.y_dim = d.y_dim;
    s.pixel_off = d.pixel_off;
  } else {
    (void)scene;
    s.view = p.scene.view;
    s.film_to_world = p.film_to_world;
    s.x_dim = p.x_dim;
    s.y_dim = p.y_dim;
    s.pixel_off = 0;
  }
  return s;
}

// Scene of global sample s (batched): largest i with sample_begin[i] <= s.
__device__ __forceinline__ uint32_t find_scene(const WfBatch &b,
                                               unsigned long long s) {
  uint32_t lo = 0, hi = b.n_scenes; // invariant: begin[lo] <= s < begin[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (b.sample_begin[mid] <= s)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <class R, bool kBatch>
__device__ __forceinline__ void wf_shade(const WfParams &p) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t group = blockIdx.x % p.n_groups;
  const bool in_range = slot < p.n_slots;
  uint4 meta = in_range ? p.st.meta[slot] : make_uint4(0, 0, 0, 0);
  bool active = in_range && (meta.w & kFlagActive);
  uint32_t scene = 0;
  if (kBatch && active)
    scene = p.st.scene[slot];

  // ---- continue the path in this slot ----
  bool push_ext = false, push_light = false, finished = false;
  V3<float> ext_o = {0.f, 0.f, 0.f}, ext_d = {0.f, 0.f, 0.f},
            light_d = {0.f, 0.f, 0.f};
  float light_target = 0.f;
  PathState<float> ns;
  uint32_t location = meta.x, sample = meta.y, counter = meta.z;
  // Fields that a step leaves unchanged are neither re-read nor re-written:
  // the light multiplier exists only while a light sample is pending, the
  // continuation direction only while the path continues, and the radiance
  // total is stored only when it changed.
  bool total_dirty = false;
  if (active) {
    SlotScene sc = slot_scene<kBatch>(p, scene);
    float4 m = p.st.mult[slot];
    const F3 lm = (meta.w & kFlagLight) ? p.st.light[slot] : F3{0.f, 0.f, 0.f};
    const F3 tot = p.st.total[slot];
    const F3 o = p.st.origin[slot];
    const F3 d = (meta.w & kFlagHasNext) ? p.st.dir[slot] : o;
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
    // The vertex record is always filled (a few register moves): a pointer
    // that is null for most lanes made the compiler keep it in scratch and
    // clear it there on every launch.
    VertexInfo<float> vi;
    vi.valid = 0u;
    const bool want_gbuffer = p.gbuffer && st.iters == 0 && sample == 0;
    finished = rendering_equation_iteration(st, rng, last, light_hit, next_hit,
                                            p.settings, sc.view, &ns, &nl, &nn, &vi);
    if (want_gbuffer && vi.valid) {
      const Material<float> &mat = sc.view.materials[vi.material];
      V3<float> alb = material_albedo(mat);
      float *g = p.gbuffer + 12 * size_t(sc.pixel_off + location);
      g[0] = vi.p.x, g[1] = vi.p.y, g[2] = vi.p.z;
      g[3] = vi.normal.x, g[4] = vi.normal.y, g[5] = vi.normal.z;
      g[6] = alb.x, g[7] = alb.y, g[8] = alb.z;
      g[9] = mat.emission.x, g[10] = mat.emission.y, g[11] = mat.emission.z;
    }
    counter = rng_counter(rng);
    total_dirty = ns.total.x != tot.x || ns.total.y != tot.y || ns.total.z != tot.z;
    if (finished) {
      float *out = p.out_samples +
                   3 * (size_t(sc.pixel_off + location) * p.rng.samples_per + sample);
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
  // Block-aggregated counters: one atomic per block on the group's sample
  // counter and one per queue (wave ballots + an LDS prefix over the block's
  // 4 waves).
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
    s_sample_base =
        tot ? atomicAdd(reinterpret_cast<unsigned long long *>(
                            group_counter(p.group_counters, kGroupSample, group)),
                        (unsigned long long)tot)
            : 0ull;
  }
  __syncthreads();
  if (want_new) {
    unsigned long long before = 0;
    for (uint32_t w = 0; w < wave; ++w)
      before += s_counts[0][w];
    const unsigned long long i = s_sample_base + before + prefix;
    if (i < group_sample_count(p, group)) {
      const unsigned long long s = group_sample(p, group, i);
      unsigned long long local = s;
      if constexpr (kBatch) {
        scene = find_scene(p.batch, s);
        local = s - p.batch.sample_begin[scene];
      }
      SlotScene sc = slot_scene<kBatch>(p, scene);
      location = uint32_t(local / p.rng.samples_per);
      sample = uint32_t(local % p.rng.samples_per);
      R rng = make_rng<R>(p.rng, location, sample, rng_start<R>(p.rng, location));
      float xo = rng.next();
      float yo = rng.next();
      uint32_t px = location % sc.x_dim, py = location / sc.x_dim;
      Ray<float> cam = camera_ray(sc.film_to_world, float(px) + xo,
                                  float(py) + yo, sc.x_dim, sc.y_dim);
      counter = rng_counter(rng);
      ns = initial_path_state<float>();
      total_dirty = true;
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
    if constexpr (kBatch)
      p.st.scene[slot] = scene;
    p.st.mult[slot] = make_float4(ns.ray_mult.x, ns.ray_mult.y, ns.ray_mult.z,
                                  ns.light_target);
    if (push_light)
      p.st.light[slot] = F3{ns.light_mult.x, ns.light_mult.y, ns.light_mult.z};
    if (total_dirty)
      p.st.total[slot] = F3{ns.total.x, ns.total.y, ns.total.z};
    p.st.origin[slot] = F3{ext_o.x, ext_o.y, ext_o.z};
    if (push_ext)
      p.st.dir[slot] = F3{ext_d.x, ext_d.y, ext_d.z};
    if (push_light)
      p.st.light_dir[slot] =
          make_float4(light_d.x, light_d.y, light_d.z, light_target);
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
    s_queue_base[0] = te ? atomicAdd(group_counter(p.group_counters, kGroupExt, group), te) : 0u;
    s_queue_base[1] =
        ts ? atomicAdd(group_counter(p.group_counters, kGroupShadow, group), ts) : 0u;
  }
  __syncthreads();
  uint32_t be = s_queue_base[0], bs = s_queue_base[1];
  for (uint32_t w = 0; w < wave; ++w) {
    be += s_counts[1][w];
    bs += s_counts[2][w];
  }
  const size_t seg = size_t(group) * p.q.group_cap;
  if (push_ext)
    p.q.ext[seg + be + pe] = slot;
  if (push_light)
    p.q.shadow[seg + bs + ps] = slot;
}

// All shade kernels require blockDim.x == 256 (block-level aggregation).
// Batches run it at 8 waves per SIMD (62 VGPRs and 80 B of spills instead of
// 74 VGPRs at 6 waves; 1024 generated scenes: render 132.0 -> 128.3 ms, and
// 7 waves measured 133.5). A single render asks for at least 6, which the
// compiler's own allocation (68 VGPRs, 7 waves, no spills) already meets;
// asking 8 there spilled more and measured slower.
template <class R, bool kBatch>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kBatch ? 8 : 6))) void
k_wf_shade(WfParams p) {
  wf_shade<R, kBatch>(p);
}

hipError_t launch_wf_shade(const WfParams &p, bool batched, hipStream_t s) {
  uint32_t blocks = (p.n_slots + 255) / 256;
  // groups and sample shares are laid out for exactly this grid
  if (blocks != p.n_shade_blocks || p.n_groups == 0 || p.n_groups > kQueueGroups ||
      p.n_groups > blocks || p.group_base + p.n_groups > p.n_groups_total)
    return hipErrorInvalidValue;
  bool sobol = p.rng.type == kRngSobel;
  if (batched) {
    if (sobol)
      hipLaunchKernelGGL((k_wf_shade<TableRng, true>), dim3(blocks), dim3(256), 0, s, p);
    else
      hipLaunchKernelGGL((k_wf_shade<HashRng, true>), dim3(blocks), dim3(256), 0, s, p);
  } else {
    if (sobol)
      hipLaunchKernelGGL((k_wf_shade<TableRng, false>), dim3(blocks), dim3(256), 0, s, p);
    else
      hipLaunchKernelGGL((k_wf_shade<HashRng, false>), dim3(blocks), dim3(256), 0, s, p);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Persistent traversal. Static schedule: wave w takes one contiguous range of
// the concatenated queues (no same-address atomics).
// ---------------------------------------------------------------------------
// Lanes are refilled once at least this many are idle (or all are): a refill
// runs the ray-setup code for the whole wave, so batching it pays (measured
// 436 / 446 / 446 / 442 Msamples/s for 1 / 8 / 16 / 32 on the 82k-triangle
// scene).
constexpr uint32_t kRefillMinIdle = 16;


// Each wave owns one contiguous range of the queue (sharing the second half
// of the queue dynamically instead, in chunks of 64 from 16 pools, measured
// 716 vs 762 Msamples/s on the 82k-triangle scene, and all of it 675: the
// contiguous ranges keep neighbouring pixels' rays in one wave) and keeps its
// 64 lanes busy: a lane whose ray finished takes the next entry of the range (ballot +
// mbcnt, no atomics) while the other lanes continue, so a long ray delays
// only itself. Lanes step one node or leaf at a time: holding leaves until
// the whole wave has one (as trace_closest does) measured 438 vs 456
// Msamples/s here, since it delays the refill. Each wave iteration runs one
// kind of step, chosen by a vote of the lanes (TraceState::step_voted): with
// both kinds run whenever the lanes were mixed, 29% of the issued lanes did
// work (SQ_THREAD_CYCLES_VALU / 64 SQ_INSTS_VALU); voting raised that to 38%
// and the 82k-triangle workload from 680 to 755-760 Msamples/s. Vote weights
// (inner if w * inner lanes >= 4 * leaf lanes) 2 / 3 / 4 / 6 / 8 / 16: 746 /
// 760 / 754 / 749 / 728 / 682; one triangle per leaf step instead of the
// whole leaf: 748. With both queues in one launch and balanced shares,
// 2 / 3 / 4 / 5 / 6: 881-883 / 895-898 / 902-903 / 899-900 / 898-899 (the
// batched scenes: no difference between 3 and 4); refill thresholds 8 / 16 /
// 24 / 32: 884 / 895-898 / 899-900 / 887.
constexpr uint32_t kTraceVote = 4;
// Both queues of an iteration in one launch (the extension segments, then
// the shadow segments, concatenated); a lane's entry decides its ray kind (a shadow ray carries its target window and stops at the first
// occluder). One launch instead of an extension launch followed by a shadow
// launch measured 757 -> 856 Msamples/s on the 82k-triangle scene (C3 fresh
// scenes 4,671 -> 5,056 scenes/s): one tail per iteration instead of two, and
// waves that finish their extension rays continue with shadow rays.
template <bool kBatch>
__device__ __forceinline__ void wf_trace(const WfParams &p) {
  __shared__ int lds_stack[kLdsStackInts * 4];
  __shared__ uint32_t s_prefix[2 * kQueueGroups + 1];
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (uint32_t g = 0; g < 2 * kQueueGroups; ++g) {
      s_prefix[g] = sum;
      sum += *group_counter(p.group_counters, g < kQueueGroups ? kGroupExt : kGroupShadow,
                            g % kQueueGroups);
    }
    s_prefix[2 * kQueueGroups] = sum;
    if (blockIdx.x == 0) {
      p.counters[kExtCount] = s_prefix[kQueueGroups];
      p.counters[kShadowCount] = sum - s_prefix[kQueueGroups];
    }
  }
  __syncthreads();
  // Each wave takes an equal share of both queues (a contiguous range of
  // each): its extension rays, then its shadow rays, as local entries
  // [0, e_len + s_len); local entry j maps to entry v of the concatenated
  // segments. Splitting the concatenated queue evenly instead left the waves
  // of the extension part with more work (extension rays cost more than
  // shadow rays): 859 vs 886 Msamples/s on the 82k-triangle scene.
  const uint32_t n_ext = s_prefix[kQueueGroups];
  const uint32_t n_shadow = s_prefix[2 * kQueueGroups] - n_ext;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t wave = gtid >> 6, n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t per_e = (n_ext + n_waves - 1) / n_waves;
  const uint32_t per_s = (n_shadow + n_waves - 1) / n_waves;
  const uint32_t e0 = min(wave * per_e, n_ext), e_len = min(e0 + per_e, n_ext) - e0;
  const uint32_t s0 = min(wave * per_s, n_shadow), s_len = min(s0 + per_s, n_shadow) - s0;
  uint32_t next = 0;
  const uint32_t end = e_len + s_len;
  if (next >= end)
    return; // whole wave
  LdsStack stack{lds_stack_column(lds_stack), p.stack_overflow + gtid, p.overflow_stride};
  TraceState ts;
  const float4 *nodes = p.scene.nodes, *tris = p.scene.tris;
  uint32_t slot = 0;
  bool has = false, shadow = false;
  for (;;) {
    const uint64_t idle = __ballot(!has);
    const uint32_t n_idle = uint32_t(__popcll(idle));
    if (next < end && (n_idle >= kRefillMinIdle || n_idle == 64u)) {
      const uint32_t pre = __builtin_amdgcn_mbcnt_hi(
          uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
      if (!has && next + pre < end) {
        const uint32_t j = next + pre;
        const uint32_t v = j < e_len ? e0 + j : n_ext + s0 + (j - e_len);
        uint32_t g = 0;
        for (uint32_t step = kQueueGroups; step; step >>= 1)
          if (s_prefix[g + step] <= v)
            g += step;
        shadow = g >= kQueueGroups;
        const uint32_t *q = shadow ? p.q.shadow : p.q.ext;
        slot = q[size_t(g % kQueueGroups) * p.q.group_cap + (v - s_prefix[g])];
        const F3 o = p.st.origin[slot];
        float4 d;
        if (shadow) {
          d = p.st.light_dir[slot];
        } else {
          const F3 e = p.st.dir[slot];
          d = make_float4(e.x, e.y, e.z, 0.f);
        }
        const Ray<float> ray = {{o.x, o.y, o.z}, {d.x, d.y, d.z}};
        bool has_nodes = true;
        if constexpr (kBatch) {
          const SceneDesc &sd = p.batch.scenes[p.st.scene[slot]];
          nodes = p.batch.nodes + kNodeFloat4 * size_t(sd.node_off);
          tris = p.batch.tris + 3 * size_t(sd.tri_off);
          has_nodes = sd.has_nodes != 0;
        }
        ts.init(ray, shadow ? d.w + 2e-6f : 3.4028235e38f, shadow ? d.w - 2e-6f : -1.0f);
        stack.sp = 0;
        if (has_nodes)
          has = true;
        else
          (shadow ? p.st.hit_light : p.st.hit_ext)[slot] = pack_hit(Hit<float>{0.f, kNoHit, 0});
      }
      next = min(next + n_idle, end);
    }
    if (__ballot(has) == 0) {
      if (next >= end)
        break;
      continue;
    }
    if (ts.step_voted(nodes, tris, stack, has, kTraceVote)) {
      (shadow ? p.st.hit_light : p.st.hit_ext)[slot] = pack_hit(ts.hit());
      has = false;
    }
  }
}

// Waves per SIMD the compiler must allow: 8 (at most 64 VGPRs) for batches
// of many small scenes (1024 generated scenes: render 133.9 -> 132.1 ms, no
// spills), the compiler's 7 (67 VGPRs) for a single render (with 8 the
// 82k-triangle scene measured 874 vs 899 Msamples/s).
template <bool kBatch>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kBatch ? 8 : 7))) void
k_wf_trace(WfParams p) {
  wf_trace<kBatch>(p);
}

hipError_t launch_wf_trace(const WfParams &p, bool batched, uint32_t grid_blocks,
                           hipStream_t s) {
  if (batched)
    hipLaunchKernelGGL((k_wf_trace<true>), dim3(grid_blocks), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((k_wf_trace<false>), dim3(grid_blocks), dim3(256), 0, s, p);
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
    const size_t first = size_t(info.x + info.y * wd.x_dim) * spp + info.start_sample;
    const uint32_t count = info.end_sample - info.start_sample;
    const float *v = samples + 3 * first;
    uint32_t s = 0;
    // four samples (three 16-byte loads) at a time when they are 16-byte
    // aligned; the additions stay in sample order
    if ((reinterpret_cast<uintptr_t>(v) & 15u) == 0) {
      const float4 *v4 = reinterpret_cast<const float4 *>(v);
      for (; s + 4 <= count; s += 4, v4 += 3) {
        const float4 a = v4[0], b = v4[1], c = v4[2];
        total += V3<float>{a.x, a.y, a.z};
        total += V3<float>{a.w, b.x, b.y};
        total += V3<float>{b.z, b.w, c.x};
        total += V3<float>{c.y, c.z, c.w};
      }
    }
    for (; s < count; ++s)
      total += V3<float>{v[3 * s], v[3 * s + 1], v[3 * s + 2]};
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
      // the writer's own groups w .. w + ng - 1, summed in place (a private
      // copy indexed at run time would live in scratch)
      const uint32_t ng = sbs / 64u;
      for (uint32_t off = ng / 2; off >= 1; off /= 2)
        for (uint32_t i = 0; i < off; ++i)
          for (int c = 0; c < 3; ++c)
            partial[w + i][c] = partial[w + i][c] + partial[w + i + off][c];
      total = {partial[w][0], partial[w][1], partial[w][2]};
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

__global__ void k_sum_samples_f64(const float *__restrict__ samples, uint32_t n_pixels,
                                  uint32_t spp, double *__restrict__ out) {
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= n_pixels)
    return;
  double r = 0.0, g = 0.0, b = 0.0;
  const float *v = samples + 3 * size_t(px) * spp;
  for (uint32_t s = 0; s < spp; ++s) {
    r += double(v[3 * s]);
    g += double(v[3 * s + 1]);
    b += double(v[3 * s + 2]);
  }
  out[3 * size_t(px)] = r / double(spp);
  out[3 * size_t(px) + 1] = g / double(spp);
  out[3 * size_t(px) + 2] = b / double(spp);
}

hipError_t launch_sum_samples_f64(const float *samples, uint32_t n_pixels,
                                  uint32_t samples_per, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_samples_f64, dim3((n_pixels + 255) / 256), dim3(256), 0, s,
                     samples, n_pixels, samples_per, out);
  return hipGetLastError();
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
