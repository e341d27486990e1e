// Sobol table, bulk closest-hit and the mega-kernel integrator.
#include "kernels.h"

#include "../core/camera.h"
#include "../core/rng.h"
#include "trace.h"
#include "wave.h"

namespace pt {
namespace gpu {

// ---------------------------------------------------------------------------
// Sobol table: one wave per (dimension, 64 consecutive samples). Lane 0 of
// the wave evaluates its point directly; every other lane contributes the
// Gray-code step v[ctz(s)], and an inclusive XOR scan across the wave (DPP)
// turns the steps into the 64 points. Dimension-major output (TableRng).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sobol_table(const uint32_t *__restrict__ dirs,
                                                     uint32_t dim_bound,
                                                     uint32_t count,
                                                     float *__restrict__ out) {
  uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint32_t lane = threadIdx.x & 63u;
  uint32_t dim = wave % dim_bound;
  uint32_t chunk = wave / dim_bound;
  uint32_t n_chunks = (count + 63u) / 64u;
  if (chunk >= n_chunks)
    return; // whole wave exits together
  const uint32_t *d = dirs + 32u * dim;
  uint32_t s0 = chunk * 64u;
  uint32_t s = s0 + lane;
  uint32_t delta = lane == 0 ? sobol_u32(d, s0) : d[__builtin_ctz(s)];
  uint32_t x = wave_xor_scan(delta);
  if (s < count)
    out[size_t(dim) * count + s] = u32_to_unit_float(x ^ sobol_scramble(dim));
}

hipError_t launch_sobol_table(const uint32_t *dirs, uint32_t dim_bound,
                              uint32_t count, float *out, hipStream_t s) {
  uint64_t waves = uint64_t(dim_bound) * ((count + 63u) / 64u);
  uint32_t blocks = uint32_t((waves + 3) / 4);
  if (blocks == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_sobol_table, dim3(blocks), dim3(256), 0, s, dirs,
                     dim_bound, count, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Bulk closest hit for a ray array (parity tests, tools).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_intersect(DeviceScene sc,
                                                   const float *__restrict__ rays,
                                                   uint32_t n, float *t,
                                                   uint32_t *idx, uint32_t *back) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float *r = rays + 6 * size_t(i);
  Ray<float> ray = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
  Hit<float> h = {0.f, kNoHit, 0};
  if (sc.n_nodes != 0) {
    PrivateStack stack;
    h = trace_closest(sc.nodes, sc.tris, ray, 3.4028235e38f, -1.0f, stack);
  }
  t[i] = h.t;
  idx[i] = h.idx;
  back[i] = h.back;
}

hipError_t launch_intersect(const DeviceScene &sc, const float *rays,
                            uint32_t n, float *t, uint32_t *idx, uint32_t *back,
                            hipStream_t s) {
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_intersect, dim3((n + 255) / 256), dim3(256), 0, s, sc,
                     rays, n, t, idx, back);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mega-kernel: every thread runs whole paths for its pixel and sample range,
// then the partial sums of a pixel's threads are reduced across lanes.
// reference: include/render/detail/integrate_image/mega_kernel/run_impl.h:140-206,
// include/integrate/rendering_equation.h:27-129
// ---------------------------------------------------------------------------
// kPerStep (OutputPerStep, reference rendering_equation.h:56-123): the
// radiance added by iteration k of a path goes to steps[k]; a path stops after
// n_steps iterations.
template <class R, bool kPerStep>
__device__ __forceinline__ V3<float> sample_path(const MegaParams &p, const ThreadInfo &info,
                                                 uint32_t location, uint32_t sample,
                                                 uint32_t *n_rays_traced, V3<float> *steps,
                                                 LdsStack &stack) {
  R rng;
  if constexpr (std::is_same_v<R, TableRng>) {
    rng = TableRng{p.rng.vals, sample,
                   initial_dimension(location, p.rng.initial_dim_bound),
                   p.rng.dim_bound, p.rng.samples_per};
  } else {
    rng = make_hash_rng(sample + location * p.rng.samples_per);
  }
  Ray<float> next_ray, light_ray;
  Hit<float> given = {0.f, kNoHit, 0};
  if (p.initial_rays) {
    const float *r = p.initial_rays + 6 * size_t(info.x);
    next_ray = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
    // InitialIdxAndDir: the first intersection is given, at distance 1 and
    // front-facing by convention (initial_ray_sample.h:58-75)
    if (p.initial_idx)
      given = {1.f, p.initial_idx[info.x], 0};
  } else {
    float xo = rng.next();
    float yo = rng.next();
    next_ray = camera_ray(p.film_to_world, float(info.x) + xo,
                          float(info.y) + yo, p.wd.x_dim, p.wd.y_dim);
  }
  light_ray = next_ray;
  PathState<float> st = initial_path_state<float>();
  for (uint32_t step = 0;; ++step) {
    if constexpr (kPerStep) {
      if (step >= p.n_steps)
        break;
      st.total = {0.f, 0.f, 0.f};
    }
    Hit<float> light_hit = {0.f, kNoHit, 0}, next_hit = {0.f, kNoHit, 0};
    if (st.n_light) {
      light_hit = trace_closest(p.scene.nodes, p.scene.tris, light_ray,
                                st.light_target + 2e-6f,
                                st.light_target - 2e-6f, stack);
      ++*n_rays_traced;
    }
    if (st.has_next) {
      if (step == 0 && p.initial_idx) {
        next_hit = given;
      } else {
        next_hit = trace_closest(p.scene.nodes, p.scene.tris, next_ray,
                                 3.4028235e38f, -1.0f, stack);
        ++*n_rays_traced;
      }
    }
    PathState<float> ns;
    Ray<float> nl = light_ray, nn = next_ray;
    bool done = rendering_equation_iteration(st, rng, next_ray, light_hit,
                                             next_hit, p.settings,
                                             p.scene.view, &ns, &nl, &nn);
    if constexpr (kPerStep)
      steps[step] += ns.total;
    if (done)
      return ns.total;
    st = ns;
    light_ray = nl;
    next_ray = nn;
  }
  return {0.f, 0.f, 0.f};
}

// The thread's samples in order; the sum adds each finished path's radiance
// in sample order (the reference's per-thread float_rgb_total).
template <class R, bool kPerStep>
__device__ __forceinline__ V3<float>
integrate_samples(const MegaParams &p, const ThreadInfo &info, uint32_t location,
                  uint32_t *rays_traced, V3<float> *steps, LdsStack &stack) {
  V3<float> total = {0.f, 0.f, 0.f};
  uint32_t n_rays_traced = 0;
  for (uint32_t sample = info.start_sample; sample < info.end_sample; ++sample) {
    const V3<float> v = sample_path<R, kPerStep>(p, info, location, sample, &n_rays_traced,
                                                 steps, stack);
    if constexpr (!kPerStep)
      total += v;
  }
  *rays_traced = n_rays_traced;
  return total;
}

// Per-pixel reduction over the sample block (lanes of one pixel are
// consecutive; groups larger than a wave go through LDS), then the store of
// the pixel's partial (divided by spp when there is one sample block).
__device__ __forceinline__ void reduce_store(const MegaParams &p, const ThreadInfo &info,
                                             uint32_t block, V3<float> total,
                                             float *out) {
  uint32_t sbs = p.wd.sample_block_size;
  uint32_t group = sbs < 64u ? sbs : 64u;
  total.x = group_reduce_sum(total.x, group);
  total.y = group_reduce_sum(total.y, group);
  total.z = group_reduce_sum(total.z, group);
  bool writer = (threadIdx.x % sbs) == 0 && !info.exit;
  if (sbs > 64u) {
    __shared__ float partial[16][3];
    uint32_t w = threadIdx.x >> 6;
    __syncthreads(); // partial may still be read by a previous call
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
    if (p.divide_by_spp) {
      float spp = float(p.rng.samples_per);
      total = {total.x / spp, total.y / spp, total.z / spp};
    }
    size_t idx = size_t(p.wd.sample_block_idx(block)) +
                 size_t(p.wd.num_sample_blocks) *
                     (size_t(info.x) + size_t(info.y) * p.wd.x_dim);
    out[3 * idx + 0] = total.x;
    out[3 * idx + 1] = total.y;
    out[3 * idx + 2] = total.z;
  }
}

// Rays traced by the block, added to the render's counter with one device
// atomic per block (atomics on one address serialise).
__device__ __forceinline__ void count_rays(const MegaParams &p, uint32_t traced) {
  if (!p.ray_counter)
    return;
  __shared__ uint32_t block_rays;
  if (threadIdx.x == 0)
    block_rays = 0;
  __syncthreads();
  uint32_t lane_total = traced;
  for (int off = 32; off >= 1; off >>= 1)
    lane_total += __shfl_xor(lane_total, off, 64);
  if ((threadIdx.x & 63u) == 0)
    atomicAdd(&block_rays, lane_total);
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(p.ray_counter, (unsigned long long)block_rays);
}

// Blocks of up to 1024 threads (16 waves); dynamic LDS holds the traversal
// stacks (kLdsStackInts per wave), deeper entries go to a private array.
// One block of the mega-kernel grid of `p` (block index within its work
// division); `mk_stack` is the block's dynamic LDS (kLdsStackInts per wave).
template <bool kPerStep>
__device__ __forceinline__ void megakernel_block(const MegaParams &p, uint32_t block,
                                                 int *mk_stack) {
  int spill[kOverflowDepth];
  LdsStack stack{lds_stack_column(mk_stack), spill, 1};
  ThreadInfo info = p.wd.get_thread_info(block, threadIdx.x);
  V3<float> total = {0.f, 0.f, 0.f};
  V3<float> steps[kPerStep ? kMaxOutputSteps : 1];
  if constexpr (kPerStep)
    for (uint32_t i = 0; i < kMaxOutputSteps; ++i)
      steps[i] = {0.f, 0.f, 0.f};
  uint32_t traced = 0;
  if (!info.exit) {
    uint32_t location = info.x + info.y * p.wd.x_dim;
    if (p.rng.type == kRngSobel)
      total = integrate_samples<TableRng, kPerStep>(p, info, location, &traced, steps, stack);
    else
      total = integrate_samples<HashRng, kPerStep>(p, info, location, &traced, steps, stack);
  }
  if constexpr (kPerStep) {
    const size_t image = size_t(p.wd.num_sample_blocks) * p.wd.x_dim * p.wd.y_dim * 3;
    for (uint32_t i = 0; i < p.n_steps; ++i)
      reduce_store(p, info, block, steps[i], p.out_rgb + i * image);
  } else {
    reduce_store(p, info, block, total, p.out_rgb);
  }
  count_rays(p, traced);
}

template <bool kPerStep>
__global__ __launch_bounds__(1024) void k_megakernel(MegaParams p,
                                                     uint32_t start_block) {
  extern __shared__ int mk_stack[];
  megakernel_block<kPerStep>(p, blockIdx.x + start_block, mk_stack);
}

// Many ray-sample-spec renders in one grid: the blocks of item i are
// [block_begin[i], block_begin[i + 1]); each block runs the item's own work
// division over its scene (offsets into the batch's concatenated arrays), so
// every item's values equal its render alone.
template <bool kPerStep>
__global__ __launch_bounds__(1024) void k_megakernel_batch(MegaBatchParams bp,
                                                           uint32_t start_block) {
  extern __shared__ int mk_stack[];
  const uint32_t block = blockIdx.x + start_block;
  uint32_t lo = 0, hi = bp.n_items; // block_begin[lo] <= block < block_begin[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (bp.block_begin[mid] <= block)
      lo = mid;
    else
      hi = mid;
  }
  const MegaBatchItem &it = bp.items[lo];
  MegaParams p{};
  p.scene.nodes = bp.nodes + kNodeFloat4 * size_t(it.node_off);
  p.scene.tris = bp.tris + 3 * size_t(it.tri_off);
  p.scene.view.tris = bp.shade + it.shade_off;
  p.scene.view.materials = bp.materials + it.mat_off;
  p.scene.view.light_tris = bp.light_tris + it.light_off;
  p.scene.view.light_cdf = bp.light_cdf + it.light_off;
  p.scene.view.n_lights = it.n_lights;
  p.settings = bp.settings;
  p.wd = it.wd;
  p.rng = bp.rng;
  p.initial_rays = bp.rays + 6 * size_t(it.ray_off);
  p.initial_idx = bp.initial_idx ? bp.initial_idx + it.ray_off : nullptr;
  p.out_rgb = bp.out + it.out_off;
  p.n_steps = bp.n_steps;
  p.divide_by_spp = it.wd.num_sample_blocks == 1 ? 1u : 0u;
  p.ray_counter = bp.ray_counter;
  megakernel_block<kPerStep>(p, block - bp.block_begin[lo], mk_stack);
}

hipError_t launch_megakernel_batch(const MegaBatchParams &bp, uint32_t block_size,
                                   uint32_t total_blocks, uint32_t max_blocks_per_launch,
                                   hipStream_t s) {
  if (block_size == 0 || block_size > 1024 || block_size % 64 != 0)
    return hipErrorInvalidConfiguration;
  const size_t lds = size_t(block_size / 64) * kLdsStackInts * sizeof(int);
  const uint32_t chunk = max_blocks_per_launch ? max_blocks_per_launch : total_blocks;
  for (uint32_t start = 0; start < total_blocks; start += chunk) {
    const uint32_t n = total_blocks - start < chunk ? total_blocks - start : chunk;
    if (bp.n_steps)
      hipLaunchKernelGGL(k_megakernel_batch<true>, dim3(n), dim3(block_size), lds, s, bp,
                         start);
    else
      hipLaunchKernelGGL(k_megakernel_batch<false>, dim3(n), dim3(block_size), lds, s, bp,
                         start);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

// Sample-parallel mega-kernel. The work division gives a thread a pixel and a
// run of consecutive samples, and the one-lane-per-thread kernel above leaves
// few waves per SIMD when there are few threads (65,536 pixels x 2 threads
// for 256 x 256 at 16 spp: 2 waves per SIMD). Here the block's lanes share
// the samples of its kSplitThreads threads (kSplit lanes per thread, samples
// handed out as lanes finish), the path radiances go to LDS by (thread,
// sample), and one lane per thread then adds them in sample order
// and the pixel's threads are reduced on consecutive lanes of wave 0 exactly
// as above, so the image is the same bit for bit. Used when the pixel's
// threads fit one wave (sample_block_size <= 64) and there is no per-step
// output.
constexpr uint32_t kSplit = 4;
constexpr uint32_t kSplitBlock = 256;                  // lanes per block
constexpr uint32_t kSplitThreads = kSplitBlock / kSplit; // work-division threads per block

// Six waves per SIMD (80 VGPRs, a few spills) measured 3-5% faster than the
// compiler's four (121 VGPRs) on the 36-triangle box.
__global__ __launch_bounds__(kSplitBlock) __attribute__((amdgpu_waves_per_eu(6, 6))) void
k_megakernel_split(MegaParams p,
                                                                  uint32_t start_block,
                                                                  uint32_t max_spt) {
  extern __shared__ int mk_lds[];
  // traversal stacks (kLdsStackInts per wave), then [kSplitThreads][max_spt] radiances
  int spill[kOverflowDepth];
  LdsStack stack{lds_stack_column(mk_lds), spill, 1};
  float *rad = reinterpret_cast<float *>(mk_lds + (kSplitBlock / 64) * kLdsStackInts);
  __shared__ uint4 s_info[kSplitThreads]; // x, y, first sample, end sample (= first: none)
  __shared__ uint32_t s_next;
  const uint32_t parts = p.wd.block_size / kSplitThreads;
  const uint32_t block = start_block + blockIdx.x / parts;
  const uint32_t first = (blockIdx.x % parts) * kSplitThreads;
  if (threadIdx.x < kSplitThreads) {
    const ThreadInfo ti = p.wd.get_thread_info(block, first + threadIdx.x);
    s_info[threadIdx.x] = ti.exit ? make_uint4(0, 0, 0, 0)
                                  : make_uint4(ti.x, ti.y, ti.start_sample, ti.end_sample);
  }
  if (threadIdx.x == 0)
    s_next = kSplitBlock;
  __syncthreads();
  // Items are (thread, sample offset) pairs, round r of kSplitBlock items
  // holding offsets kSplit r .. kSplit r + kSplit - 1 of every thread: lane l
  // starts with item l (offset l % kSplit of thread l / kSplit), then takes
  // the block's next unclaimed item whenever its path ends, so a lane with
  // long paths no longer holds its wave while the others idle (36-triangle
  // box, 256 x 256, 16 spp: 2,245-2,257 -> 2,268-2,271 Msamples/s against a
  // fixed stride of samples per lane). Radiances go to their (thread, offset)
  // slot, so the sums below are unchanged.
  const uint32_t n_items = kSplitBlock * ((max_spt + kSplit - 1) / kSplit);
  uint32_t traced = 0;
  for (uint32_t k = threadIdx.x; k < n_items; k = atomicAdd(&s_next, 1u)) {
    const uint32_t w = k % kSplitBlock, t = w / kSplit;
    const uint32_t o = w % kSplit + kSplit * (k / kSplitBlock);
    const uint4 ti = s_info[t];
    if (ti.z + o >= ti.w)
      continue; // the thread has fewer samples (or none)
    const ThreadInfo info = {ti.x, ti.y, ti.z, ti.w, false};
    const uint32_t location = info.x + info.y * p.wd.x_dim;
    const uint32_t sample = info.start_sample + o;
    const V3<float> v =
        p.rng.type == kRngSobel
            ? sample_path<TableRng, false>(p, info, location, sample, &traced, nullptr, stack)
            : sample_path<HashRng, false>(p, info, location, sample, &traced, nullptr, stack);
    float *d = rad + 3 * (size_t(t) * max_spt + o);
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // wave 0: lane i is thread first + i (kSplitThreads == 64)
    const ThreadInfo own = p.wd.get_thread_info(block, first + threadIdx.x);
    V3<float> total = {0.f, 0.f, 0.f};
    if (!own.exit) {
      const float *src = rad + 3 * size_t(threadIdx.x) * max_spt;
      for (uint32_t k = 0; k < own.end_sample - own.start_sample; ++k)
        total += V3<float>{src[3 * k], src[3 * k + 1], src[3 * k + 2]};
    }
    const uint32_t sbs = p.wd.sample_block_size;
    total.x = group_reduce_sum(total.x, sbs);
    total.y = group_reduce_sum(total.y, sbs);
    total.z = group_reduce_sum(total.z, sbs);
    if ((threadIdx.x % sbs) == 0 && !own.exit) {
      if (p.divide_by_spp) {
        const float spp = float(p.rng.samples_per);
        total = {total.x / spp, total.y / spp, total.z / spp};
      }
      const size_t idx = size_t(p.wd.sample_block_idx(block)) +
                         size_t(p.wd.num_sample_blocks) *
                             (size_t(own.x) + size_t(own.y) * p.wd.x_dim);
      p.out_rgb[3 * idx + 0] = total.x;
      p.out_rgb[3 * idx + 1] = total.y;
      p.out_rgb[3 * idx + 2] = total.z;
    }
  }
  count_rays(p, traced);
}

hipError_t launch_megakernel(const MegaParams &p, uint32_t start_block,
                             uint32_t end_block, bool split, hipStream_t s) {
  if (end_block <= start_block)
    return hipSuccess;
  if (p.wd.block_size == 0 || p.wd.block_size > 1024 || p.wd.block_size % 64 != 0)
    return hipErrorInvalidConfiguration;
  int dev = 0, lds_max = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess)
    return hipErrorInvalidDevice;
  const uint32_t max_spt =
      p.wd.base_samples_per_thread + (p.wd.n_threads_per_unit_extra ? 1u : 0u);
  const size_t split_lds = size_t(kSplitBlock / 64) * kLdsStackInts * sizeof(int) +
                           size_t(kSplitThreads) * max_spt * 3 * sizeof(float);
  // static LDS of the split kernel: the threads' sample ranges and the item
  // counter, plus slack
  const size_t split_static = sizeof(uint4) * kSplitThreads + 256;
  if (split && !p.n_steps && p.wd.sample_block_size <= 64 && max_spt >= 2 &&
      split_lds + split_static <= size_t(lds_max)) {
    const uint32_t parts = p.wd.block_size / kSplitThreads;
    hipLaunchKernelGGL(k_megakernel_split, dim3((end_block - start_block) * parts),
                       dim3(kSplitBlock), split_lds, s, p, start_block, max_spt);
    return hipGetLastError();
  }
  const size_t lds = size_t(p.wd.block_size / 64) * kLdsStackInts * sizeof(int);
  if (lds + 256 > size_t(lds_max)) // 256: the reduction's static LDS
    return hipErrorInvalidConfiguration;
  if (p.n_steps)
    hipLaunchKernelGGL(k_megakernel<true>, dim3(end_block - start_block),
                       dim3(p.wd.block_size), lds, s, p, start_block);
  else
    hipLaunchKernelGGL(k_megakernel<false>, dim3(end_block - start_block),
                       dim3(p.wd.block_size), lds, s, p, start_block);
  return hipGetLastError();
}

// The mega-kernel's per-pixel reduction applied to given per-sample values
// (the reduction of the reference's reduce test, tests/src/reduce.cu:49-158):
// each work-division thread sums in[sample + location * samples_per] over its
// samples in order, then reduce_store combines a location's threads.
__global__ __launch_bounds__(1024) void k_reduce_samples(MegaParams p,
                                                         const float *__restrict__ in,
                                                         uint32_t start_block) {
  const uint32_t block = blockIdx.x + start_block;
  const ThreadInfo info = p.wd.get_thread_info(block, threadIdx.x);
  V3<float> total = {0.f, 0.f, 0.f};
  if (!info.exit) {
    const size_t base = (size_t(info.x) + size_t(info.y) * p.wd.x_dim) * p.rng.samples_per;
    for (uint32_t i = info.start_sample; i < info.end_sample; ++i)
      total.x += in[base + i];
  }
  reduce_store(p, info, block, total, p.out_rgb);
}

hipError_t launch_reduce_samples(const WorkDivision &wd, const float *in, uint32_t samples_per,
                                 float *out_rgb, hipStream_t s) {
  if (wd.block_size == 0 || wd.block_size > 1024 || wd.block_size % 64 != 0)
    return hipErrorInvalidConfiguration;
  MegaParams p{};
  p.wd = wd;
  p.rng.samples_per = samples_per;
  p.out_rgb = out_rgb;
  p.divide_by_spp = false;
  const uint32_t total = wd.total_num_blocks();
  if (total == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_reduce_samples, dim3(total), dim3(wd.block_size), 0, s, p, in, 0u);
  return hipGetLastError();
}

__global__ void k_reduce_sample_blocks(const float *__restrict__ in, uint32_t nsb,
                                       uint32_t n_pixels, uint32_t spp,
                                       float *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pixels)
    return;
  float r = 0.f, g = 0.f, b = 0.f;
  for (uint32_t k = 0; k < nsb; ++k) {
    const float *v = in + 3 * (size_t(i) * nsb + k);
    r = r + v[0];
    g = g + v[1];
    b = b + v[2];
  }
  float f = float(spp);
  out[3 * size_t(i) + 0] = r / f;
  out[3 * size_t(i) + 1] = g / f;
  out[3 * size_t(i) + 2] = b / f;
}

hipError_t launch_reduce_sample_blocks(const float *in, uint32_t nsb,
                                       uint32_t n_pixels, uint32_t samples_per,
                                       float *out, hipStream_t s) {
  if (n_pixels == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_reduce_sample_blocks, dim3((n_pixels + 255) / 256),
                     dim3(256), 0, s, in, nsb, n_pixels, samples_per, out);
  return hipGetLastError();
}

__global__ void k_to_bgra(const float *__restrict__ rgb, uint32_t n,
                          uint32_t *__restrict__ bgra) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  bgra[i] = float_rgb_to_bgra32(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
}

hipError_t launch_to_bgra(const float *rgb, uint32_t n, uint32_t *bgra,
                          hipStream_t s) {
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_to_bgra, dim3((n + 255) / 256), dim3(256), 0, s, rgb, n,
                     bgra);
  return hipGetLastError();
}

} // namespace gpu
} // namespace pt
