// Path-replay adjoint (reverse mode) of the mega-kernel estimator.
//
// Each thread owns a run of samples of one pixel. For every sample it replays
// the path from its Sobol coordinates twice: the first replay yields the
// path's radiance L; the second walks the same vertices again and
// back-propagates dI (the loss gradient of the pixel):
//   - an emission term  mult * Le[m]  adds  dI * mult  to dLe[m];
//   - a vertex with material m scales every later term by albedo[m] exactly
//     once (single-lobe BSDFs: sample weight and NEE evaluation are linear in
//     the albedo color), so it adds  dI * (L - radiance so far) / albedo[m]
//     to dAlbedo[m] (channels with zero albedo are skipped).
// Nothing is stored per vertex. A block is one wave; its gradients accumulate
// in one LDS row of 6 x n_mats floats with LDS float atomics (ds_add_f32;
// all of a row's updates come from its one wave, in program order), and the
// rows are summed over blocks in FP64. A row per lane instead (no atomics)
// took 6 x n_mats x 65 floats per wave, 14 KB for 9 materials, which held
// the kernel to 2 waves per SIMD; the row per wave leaves the VGPR limit (4).
//
// Single pass (paths of at most kAdjointMaxVertices vertices, e.g. NIters):
// dI (L - R_k) / albedo = dI L / albedo - dI R_k / albedo, so one replay adds
// the second part at each vertex (R_k is known there), keeps the vertex's
// material in LDS, and adds the first part for the kept materials once L is
// known at the end of the path. A path longer than the list sets `overflow`
// and the host reruns the two-replay kernel.
#include "kernels.h"

#include "../core/camera.h"
#include "../core/rng.h"
#include "trace.h"

namespace pt {
namespace gpu {

template <class R>
__device__ __forceinline__ R adjoint_rng(const RngParams &rp, uint32_t location,
                                         uint32_t sample) {
  if constexpr (std::is_same_v<R, TableRng>)
    return TableRng{rp.vals, sample,
                    initial_dimension(location, rp.initial_dim_bound), rp.dim_bound,
                    rp.samples_per};
  else
    return make_hash_rng(sample + location * rp.samples_per);
}

enum ReplayMode { kRadiance, kBackprop, kSinglePass };

// Add v to LDS gradient slot `slot` (shared by the lanes of the wave).
__device__ __forceinline__ void acc_add(float *slot, float v) { atomicAdd(slot, v); }

// One replay of a sample. kRadiance: returns the radiance. kBackprop:
// accumulates into acc[param] (the wave's row) given the sample's radiance.
// kSinglePass: accumulates without knowing the radiance in advance; vlist
// (this thread's LDS column, stride kAdjointBlock) keeps the vertex
// materials.
template <class R, int kMode>
__device__ __forceinline__ V3<float>
replay_sample(const AdjointParams &p, uint32_t px, uint32_t py, uint32_t sample,
              V3<float> radiance, V3<float> d_i, float *acc, LdsStack &stack,
              uint8_t *vlist = nullptr) {
  uint32_t n_vertices = 0;
  const uint32_t location = px + py * p.x_dim;
  R rng = adjoint_rng<R>(p.rng, location, sample);
  float xo = rng.next();
  float yo = rng.next();
  Ray<float> next_ray =
      camera_ray(p.film_to_world, float(px) + xo, float(py) + yo, p.x_dim, p.y_dim);
  Ray<float> light_ray = next_ray;
  PathState<float> st = initial_path_state<float>();
  const uint32_t n3 = 3 * p.n_mats;
  for (;;) {
    Hit<float> light_hit = {0.f, kNoHit, 0}, next_hit = {0.f, kNoHit, 0};
    if (st.n_light) {
      light_hit = trace_closest(p.scene.nodes, p.scene.tris, light_ray,
                                st.light_target + 2e-6f, st.light_target - 2e-6f, stack);
    }
    if (st.has_next) {
      next_hit = trace_closest(p.scene.nodes, p.scene.tris, next_ray, 3.4028235e38f,
                               -1.0f, stack);
    }
    PathState<float> ns;
    Ray<float> nl = light_ray, nn = next_ray;
    IterTape<float> tape;
    VertexInfo<float> *no_vertex = nullptr;
    IterTape<float> *tape_ptr = kMode != kRadiance ? &tape : nullptr;
    bool done = rendering_equation_iteration(st, rng, next_ray, light_hit, next_hit,
                                             p.settings, p.scene.view, &ns, &nl, &nn,
                                             no_vertex, tape_ptr);
    if constexpr (kMode != kRadiance) {
      for (uint32_t j = 0; j < tape.n_emit; ++j) {
        float *e = acc + n3 + 3 * tape.emit_mat[j];
        acc_add(e, d_i.x * tape.emit_mult[j].x);
        acc_add(e + 1, d_i.y * tape.emit_mult[j].y);
        acc_add(e + 2, d_i.z * tape.emit_mult[j].z);
      }
      if (tape.vertex_mat != kNoMaterial) {
        V3<float> rho = material_albedo(p.scene.view.materials[tape.vertex_mat]);
        float *a = acc + 3 * tape.vertex_mat;
        if constexpr (kMode == kBackprop) {
          V3<float> rest = radiance - ns.total;
          if (rho.x != 0.f)
            acc_add(a, d_i.x * rest.x / rho.x);
          if (rho.y != 0.f)
            acc_add(a + 1, d_i.y * rest.y / rho.y);
          if (rho.z != 0.f)
            acc_add(a + 2, d_i.z * rest.z / rho.z);
        } else {
          if (rho.x != 0.f)
            acc_add(a, -(d_i.x * ns.total.x / rho.x));
          if (rho.y != 0.f)
            acc_add(a + 1, -(d_i.y * ns.total.y / rho.y));
          if (rho.z != 0.f)
            acc_add(a + 2, -(d_i.z * ns.total.z / rho.z));
          if (n_vertices < kAdjointMaxVertices)
            vlist[n_vertices * kAdjointBlock] = uint8_t(tape.vertex_mat);
          else
            *p.overflow = 1u;
          ++n_vertices;
        }
      }
    }
    if (done) {
      if constexpr (kMode == kSinglePass) {
        const V3<float> l = ns.total;
        const uint32_t nv = n_vertices < kAdjointMaxVertices ? n_vertices : kAdjointMaxVertices;
        for (uint32_t v = 0; v < nv; ++v) {
          const uint32_t m = vlist[v * kAdjointBlock];
          const V3<float> rho = material_albedo(p.scene.view.materials[m]);
          float *a = acc + 3 * m;
          if (rho.x != 0.f)
            acc_add(a, d_i.x * l.x / rho.x);
          if (rho.y != 0.f)
            acc_add(a + 1, d_i.y * l.y / rho.y);
          if (rho.z != 0.f)
            acc_add(a + 2, d_i.z * l.z / rho.z);
        }
      }
      return ns.total;
    }
    st = ns;
    light_ray = nl;
    next_ray = nn;
  }
}

template <class R, bool kSingle>
__global__ __launch_bounds__(kAdjointBlock) void k_adjoint(AdjointParams p) {
  extern __shared__ float acc_all[];
  __shared__ uint8_t vlist_all[kSingle ? kAdjointMaxVertices * kAdjointBlock : 1];
  __shared__ int stack_lds[kLdsStackInts * (kAdjointBlock / 64)];
  int spill[kOverflowDepth];
  LdsStack stack{lds_stack_column(stack_lds), spill, 1};
  const uint32_t n_params = 6 * p.n_mats;
  for (uint32_t i = threadIdx.x; i < n_params; i += kAdjointBlock)
    acc_all[i] = 0.f;
  __syncthreads();
  const uint32_t t = blockIdx.x * kAdjointBlock + threadIdx.x;
  if (t < p.n_threads) {
    const uint32_t pixel = t / p.threads_per_pixel;
    const uint32_t part = t % p.threads_per_pixel;
    const uint32_t px = pixel % p.x_dim, py = pixel / p.x_dim;
    const uint32_t spp = p.rng.samples_per;
    const uint32_t s0 = part * p.samples_per_thread;
    const uint32_t s1 = s0 + p.samples_per_thread < spp ? s0 + p.samples_per_thread : spp;
    const V3<float> d_i = {p.d_image[3 * size_t(pixel)], p.d_image[3 * size_t(pixel) + 1],
                           p.d_image[3 * size_t(pixel) + 2]};
    for (uint32_t sample = s0; sample < s1; ++sample) {
      V3<float> zero = {0.f, 0.f, 0.f};
      if constexpr (kSingle) {
        replay_sample<R, kSinglePass>(p, px, py, sample, zero, d_i, acc_all, stack,
                                      vlist_all + threadIdx.x);
      } else {
        V3<float> radiance =
            replay_sample<R, kRadiance>(p, px, py, sample, zero, zero, nullptr, stack);
        replay_sample<R, kBackprop>(p, px, py, sample, radiance, d_i, acc_all, stack);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n_params; i += kAdjointBlock)
    p.partials[size_t(blockIdx.x) * n_params + i] = acc_all[i];
}

hipError_t launch_adjoint(const AdjointParams &p, hipStream_t s) {
  const uint32_t blocks = (p.n_threads + kAdjointBlock - 1) / kAdjointBlock;
  const size_t lds = size_t(6) * p.n_mats * sizeof(float);
  const bool sobol = p.rng.type == kRngSobel;
  if (p.single_pass) {
    if (sobol)
      hipLaunchKernelGGL((k_adjoint<TableRng, true>), dim3(blocks), dim3(kAdjointBlock), lds, s, p);
    else
      hipLaunchKernelGGL((k_adjoint<HashRng, true>), dim3(blocks), dim3(kAdjointBlock), lds, s, p);
  } else {
    if (sobol)
      hipLaunchKernelGGL((k_adjoint<TableRng, false>), dim3(blocks), dim3(kAdjointBlock), lds, s,
                         p);
    else
      hipLaunchKernelGGL((k_adjoint<HashRng, false>), dim3(blocks), dim3(kAdjointBlock), lds, s,
                         p);
  }
  return hipGetLastError();
}

// Sum of the per-wave gradient rows (n_blocks rows of n_params floats) in
// two passes, both in a fixed order (deterministic). Pass 1: workgroup g sums
// the rows of its contiguous range; its threads cover floor(256 / n_params)
// whole rows per step, so each step loads consecutive floats, and thread t
// keeps parameter t % n_params of row offset t / n_params in FP64; the row
// offsets' sums are then added in order. Pass 2: one workgroup per parameter
// adds the ranges' sums (at most 256, one per thread) with a fixed-shape tree
// and scales. (Pass 2 as one thread per parameter looping over the ranges
// took 59 us: 256 dependent loads per thread.)
__global__ __launch_bounds__(256) void k_adjoint_rows(const float *__restrict__ partials,
                                                      uint32_t n_blocks, uint32_t n_params,
                                                      uint32_t rows_per_group,
                                                      double *__restrict__ group_sums) {
  __shared__ double sums[256];
  const uint32_t per_step = 256u / n_params;
  const uint32_t t = threadIdx.x;
  const uint32_t first = blockIdx.x * rows_per_group;
  const uint32_t last = min(first + rows_per_group, n_blocks);
  double sum = 0.0;
  if (t < per_step * n_params) {
    const uint32_t param = t % n_params;
    for (uint32_t r = first + t / n_params; r < last; r += per_step)
      sum += double(partials[size_t(r) * n_params + param]);
  }
  sums[t] = sum;
  __syncthreads();
  if (t < n_params) {
    double total = 0.0;
    for (uint32_t k = 0; k < per_step; ++k)
      total += sums[k * n_params + t];
    group_sums[size_t(blockIdx.x) * n_params + t] = total;
  }
}

static_assert(kAdjointFinalizeGroups <= 256, "pass 2 takes one range per thread");

__global__ __launch_bounds__(256) void k_adjoint_finalize(const double *__restrict__ group_sums,
                                                          uint32_t n_groups, uint32_t n_params,
                                                          double scale,
                                                          float *__restrict__ out) {
  __shared__ double sums[256];
  const uint32_t i = blockIdx.x, g = threadIdx.x;
  sums[g] = g < n_groups ? group_sums[size_t(g) * n_params + i] : 0.0;
  __syncthreads();
  for (uint32_t off = 128u; off >= 1u; off >>= 1) {
    if (g < off)
      sums[g] += sums[g + off];
    __syncthreads();
  }
  if (g == 0)
    out[i] = float(sums[0] * scale);
}

uint32_t adjoint_finalize_groups(uint32_t n_blocks) {
  const uint32_t rows = (n_blocks + kAdjointFinalizeGroups - 1) / kAdjointFinalizeGroups;
  return rows ? (n_blocks + rows - 1) / rows : 0;
}

hipError_t launch_adjoint_finalize(const float *partials, uint32_t n_blocks,
                                   uint32_t n_params, double scale, double *group_sums,
                                   float *out, hipStream_t s) {
  if (n_params == 0 || n_params > 256 || n_blocks == 0)
    return hipErrorInvalidValue;
  const uint32_t rows = (n_blocks + kAdjointFinalizeGroups - 1) / kAdjointFinalizeGroups;
  const uint32_t groups = adjoint_finalize_groups(n_blocks);
  hipLaunchKernelGGL(k_adjoint_rows, dim3(groups), dim3(256), 0, s, partials, n_blocks,
                     n_params, rows, group_sums);
  hipLaunchKernelGGL(k_adjoint_finalize, dim3(n_params), dim3(256), 0, s, group_sums, groups,
                     n_params, scale, out);
  return hipGetLastError();
}

// d = 2 (image - target), each difference in double and rounded once to
// float: the gradient of sum((image - target)^2).
__global__ void k_l2_gradient(const float *__restrict__ image, const float *__restrict__ target,
                              uint64_t n, float *__restrict__ d) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    d[i] = float(2.0 * (double(image[i]) - double(target[i])));
}

hipError_t launch_l2_gradient(const float *image, const float *target, uint64_t n, float *d,
                              hipStream_t s) {
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_l2_gradient, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, image,
                     target, n, d);
  return hipGetLastError();
}

} // namespace gpu
} // namespace pt
