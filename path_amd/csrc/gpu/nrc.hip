// One-bounce paths for the radiance-cache estimator (see NrcParams).
#include "kernels.h"

#include "../core/camera.h"
#include "../core/rng.h"
#include "trace.h"

namespace pt {
namespace gpu {

template <class R>
__global__ __launch_bounds__(256) void k_nrc_paths(NrcParams p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const uint32_t spp = p.rng.samples_per;
  const uint32_t location = i / spp, sample = i % spp;
  const uint32_t px = location % p.x_dim, py = location / p.x_dim;
  R rng;
  if constexpr (std::is_same_v<R, TableRng>)
    rng = TableRng{p.rng.vals, sample,
                   initial_dimension(location, p.rng.initial_dim_bound), p.rng.dim_bound,
                   p.rng.samples_per};
  else
    rng = make_hash_rng(sample + location * spp);
  const float xo = rng.next();
  const float yo = rng.next();
  Ray<float> cam = camera_ray(p.film_to_world, float(px) + xo, float(py) + yo, p.x_dim, p.y_dim);
  const SceneView<float> &sc = p.scene.view;
  V3<float> base = {0.f, 0.f, 0.f}, weight = {0.f, 0.f, 0.f};
  float feat[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  __shared__ int stack_lds[kLdsStackInts * 4];
  int spill[kOverflowDepth];
  LdsStack stack{lds_stack_column(stack_lds), spill, 1};
  Hit<float> h1 = trace_closest(p.scene.nodes, p.scene.tris, cam, 3.4028235e38f, -1.0f, stack);
  PathState<float> st = initial_path_state<float>(), ns;
  Ray<float> nl = cam, nn = cam;
  const Hit<float> none = {0.f, kNoHit, 0};
  const bool done = rendering_equation_iteration(st, rng, cam, none, h1, p.settings, sc, &ns,
                                                 &nl, &nn);
  base = ns.total;
  if (!done) {
    // light sample of the first vertex
    if (ns.n_light) {
      Hit<float> lh = trace_closest(p.scene.nodes, p.scene.tris, nl, ns.light_target + 2e-6f,
                                    ns.light_target - 2e-6f, stack);
      if (lh.idx != kNoHit && within_target(ns.light_target, lh.t) &&
          !(p.settings.back_cull_emission && lh.back))
        base += sc.materials[sc.tris[lh.idx].material].emission * ns.light_mult;
    }
    // second vertex: emission only after a discrete bounce, the rest is cached
    if (ns.has_next) {
      Hit<float> h2 = trace_closest(p.scene.nodes, p.scene.tris, nn, 3.4028235e38f, -1.0f,
                                    stack);
      if (h2.idx != kNoHit) {
        const ShadeTri<float> tri = sc.tris[h2.idx];
        const Material<float> mat = sc.materials[tri.material];
        if (ns.count_emission && !(p.settings.back_cull_emission && h2.back))
          base += ns.ray_mult * mat.emission;
        const V3<float> q = hit_point(nn, h2.t);
        const TriGeom<float> g = {tri.v0, tri.e1, tri.e2};
        const V3<float> vn[3] = {tri.n0, tri.n1, tri.n2};
        const V3<float> nrm = shading_normal(g, vn, q);
        const V3<float> alb = material_albedo(mat);
        weight = ns.ray_mult;
        feat[0] = q.x, feat[1] = q.y, feat[2] = q.z;
        feat[3] = nrm.x, feat[4] = nrm.y, feat[5] = nrm.z;
        feat[6] = alb.x, feat[7] = alb.y, feat[8] = alb.z;
      }
    }
  }
  p.base[3 * size_t(i)] = base.x;
  p.base[3 * size_t(i) + 1] = base.y;
  p.base[3 * size_t(i) + 2] = base.z;
  p.weight[3 * size_t(i)] = weight.x;
  p.weight[3 * size_t(i) + 1] = weight.y;
  p.weight[3 * size_t(i) + 2] = weight.z;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    p.features[9 * size_t(i) + k] = feat[k];
}

hipError_t launch_nrc_paths(const NrcParams &p, hipStream_t s) {
  const uint64_t n64 = uint64_t(p.x_dim) * p.y_dim * p.rng.samples_per;
  if (n64 >= (1ull << 32))
    return hipErrorInvalidValue;
  const uint32_t n = uint32_t(n64);
  if (p.rng.type == kRngSobel)
    hipLaunchKernelGGL(k_nrc_paths<TableRng>, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  else
    hipLaunchKernelGGL(k_nrc_paths<HashRng>, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// First camera hit of each pixel through the pixel centre: the MLP's
// features there, the emission seen (base) and weight 1 when something is hit.
__global__ __launch_bounds__(256) void k_nrc_first_vertex(NrcParams p, uint32_t n_pixels) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ int stack_lds[kLdsStackInts * 4];
  if (i >= n_pixels)
    return;
  const uint32_t px = i % p.x_dim, py = i / p.x_dim;
  const Ray<float> cam = camera_ray(p.film_to_world, float(px) + 0.5f, float(py) + 0.5f,
                                    p.x_dim, p.y_dim);
  int spill[kOverflowDepth];
  LdsStack stack{lds_stack_column(stack_lds), spill, 1};
  const Hit<float> h = trace_closest(p.scene.nodes, p.scene.tris, cam, 3.4028235e38f, -1.0f,
                                     stack);
  V3<float> base = {0.f, 0.f, 0.f};
  float w = 0.f;
  float feat[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (h.idx != kNoHit) {
    const SceneView<float> &sc = p.scene.view;
    const ShadeTri<float> tri = sc.tris[h.idx];
    const Material<float> mat = sc.materials[tri.material];
    if (!(p.settings.back_cull_emission && h.back))
      base = mat.emission;
    const V3<float> q = hit_point(cam, h.t);
    const TriGeom<float> g = {tri.v0, tri.e1, tri.e2};
    const V3<float> vn[3] = {tri.n0, tri.n1, tri.n2};
    const V3<float> nrm = shading_normal(g, vn, q);
    const V3<float> alb = material_albedo(mat);
    w = 1.f;
    feat[0] = q.x, feat[1] = q.y, feat[2] = q.z;
    feat[3] = nrm.x, feat[4] = nrm.y, feat[5] = nrm.z;
    feat[6] = alb.x, feat[7] = alb.y, feat[8] = alb.z;
  }
  p.base[3 * size_t(i)] = base.x;
  p.base[3 * size_t(i) + 1] = base.y;
  p.base[3 * size_t(i) + 2] = base.z;
  p.weight[3 * size_t(i)] = w;
  p.weight[3 * size_t(i) + 1] = w;
  p.weight[3 * size_t(i) + 2] = w;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    p.features[9 * size_t(i) + k] = feat[k];
}

hipError_t launch_nrc_first_vertex(const NrcParams &p, hipStream_t s) {
  const uint64_t n = uint64_t(p.x_dim) * p.y_dim;
  if (n >= (1ull << 32))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_nrc_first_vertex, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, p,
                     uint32_t(n));
  return hipGetLastError();
}

__global__ void k_nrc_combine(const float *__restrict__ base, const float *__restrict__ weight,
                              const float *__restrict__ pred, uint32_t n_pixels, uint32_t spp,
                              float *__restrict__ out) {
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= n_pixels)
    return;
  float r = 0.f, g = 0.f, b = 0.f;
  for (uint32_t s = 0; s < spp; ++s) {
    const size_t i = size_t(px) * spp + s;
    r += base[3 * i] + weight[3 * i] * fmaxf(pred[16 * i], 0.f);
    g += base[3 * i + 1] + weight[3 * i + 1] * fmaxf(pred[16 * i + 1], 0.f);
    b += base[3 * i + 2] + weight[3 * i + 2] * fmaxf(pred[16 * i + 2], 0.f);
  }
  const float inv = 1.f / float(spp);
  out[3 * size_t(px)] = r * inv;
  out[3 * size_t(px) + 1] = g * inv;
  out[3 * size_t(px) + 2] = b * inv;
}

hipError_t launch_nrc_combine(const float *base, const float *weight, const float *pred,
                              uint32_t n_pixels, uint32_t spp, float *out_rgb,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_nrc_combine, dim3((n_pixels + 255) / 256), dim3(256), 0, s, base,
                     weight, pred, n_pixels, spp, out_rgb);
  return hipGetLastError();
}

} // namespace gpu
} // namespace pt
