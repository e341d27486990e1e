// CPU checker for the GPU path tracer. TEST INFRASTRUCTURE ONLY: loaded by the
// tests, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg; nothing
// in the product links or calls it.
//
// It restates the reference's algorithm in plain sequential C++, templated on
// the scalar type (float: bit-for-bit twin of the GPU arithmetic; double: the
// FP64 oracle), reusing the per-vertex math of path_amd/csrc/core (pinned by
// its own closed-form tests):
//   - closest hit: brute force over all triangles (LoopAll,
//     include/intersect/accel/loop_all/loop_all.h:15-42) and the reference BVH
//     traversal (bvh_impl.h:12-89: DFS, left child first, leaves gathered until
//     target_objects triangles are pending) over the host SAH tree;
//   - Sobol table evaluated point by point (direct Gray-code XOR);
//   - the path loop of rendering_equation.h:27-129 with full closest hits for
//     light rays, per-thread sequential sums over the work division of
//     src/kernel/work_division.cpp and the per-pixel reduction tree;
//   - per-ray traversal counts (nodes visited, triangles tested) for the
//     algorithmic-bytes roofline.
//
// Deviations shared with the GPU path (documented in the product): equal-t
// ties go to the smaller flattened triangle index; BVH boxes are padded by
// bvh_padding() and a box is skipped only if its entry distance exceeds the
// best hit (so results do not depend on traversal order).
//
// Parity pinning: see tests/test_oracle_*.py (brute force vs BVH, reference
// KATs, Sobol closed forms, furnace and direct-lighting closed forms).
#include "../include/path_amd.h"
#include "../path_amd/csrc/core/camera.h"
#include "../path_amd/csrc/core/rng.h"
#include "../path_amd/csrc/core/shading.h"
#include "../path_amd/csrc/core/work_division.h"
#include "../path_amd/csrc/host/bvh.h"

#include <omp.h>

#include <atomic>
#include <cstring>
#include <vector>

using namespace pt;

namespace {

template <class T> struct OScene {
  std::vector<TriGeom<T>> tris;
  std::vector<ShadeTri<T>> shade;
  std::vector<Material<T>> mats;
  std::vector<TriGeom<T>> light_tris;
  std::vector<T> light_cdf;
  host::Bvh bvh;
  std::vector<V3<T>> lo, hi; // padded node boxes
  SceneView<T> view() const {
    return {shade.data(), mats.data(), light_tris.data(), light_cdf.data(),
            uint32_t(light_tris.size())};
  }
};

template <class T> TriGeom<T> load_tri(const float *p) {
  return {{T(p[0]), T(p[1]), T(p[2])},
          {T(p[3]), T(p[4]), T(p[5])},
          {T(p[6]), T(p[7]), T(p[8])}};
}

std::vector<TriGeom<float>> float_tris(const float *tris, uint32_t n) {
  std::vector<TriGeom<float>> v(n);
  for (uint32_t i = 0; i < n; ++i)
    v[i] = load_tri<float>(tris + 9 * i);
  return v;
}

template <class T>
void build_accel(OScene<T> &s, const std::vector<TriGeom<float>> &ftris) {
  s.bvh = host::build_sah_bvh(ftris);
  float pad = host::bvh_padding(ftris);
  s.lo.resize(s.bvh.nodes.size());
  s.hi.resize(s.bvh.nodes.size());
  for (size_t i = 0; i < s.bvh.nodes.size(); ++i) {
    const auto &n = s.bvh.nodes[i];
    s.lo[i] = {T(n.lo.x - pad), T(n.lo.y - pad), T(n.lo.z - pad)};
    s.hi[i] = {T(n.hi.x + pad), T(n.hi.y + pad), T(n.hi.z + pad)};
  }
}

struct Counts {
  uint64_t nodes = 0, tris = 0;
};

// Traversal counts split by ray kind: [0] continuation/camera rays,
// [1] light (shadow) rays.
struct KindCounts {
  Counts c[2];
  uint64_t rays[2] = {0, 0};
};

template <class T>
Hit<T> loop_all(const std::vector<TriGeom<T>> &tris, const Ray<T> &ray) {
  Hit<T> best{T(0), kNoHit, 0};
  for (uint32_t i = 0; i < tris.size(); ++i) {
    T t;
    uint32_t back;
    if (intersect_triangle(ray, tris[i], &t, &back) &&
        (best.idx == kNoHit || hit_before(t, i, best.t, best.idx)))
      best = {t, i, back};
  }
  return best;
}

// reference: include/intersect/accel/detail/bvh/bvh_impl.h:12-89
template <class T>
Hit<T> bvh_closest(const OScene<T> &s, const Ray<T> &ray, Counts *c) {
  Hit<T> best{T(0), kNoHit, 0};
  if (s.bvh.nodes.empty())
    return best;
  const uint32_t target_objects = 4;
  V3<T> inv = inv_direction(ray.d);
  uint32_t stack[128];
  uint32_t sp = 0;
  stack[sp++] = 0;
  uint32_t objects[64];
  uint32_t n_obj = 0;
  while (sp != 0) {
    while (n_obj < target_objects && sp != 0) {
      uint32_t ni = stack[--sp];
      const host::BvhNode &node = s.bvh.nodes[ni];
      if (c)
        ++c->nodes;
      T tmin, tmax;
      bool hit = slab_test(s.lo[ni], s.hi[ni], ray.o, inv, &tmin, &tmax);
      if (hit && (best.idx == kNoHit || !(best.t < tmin))) {
        if (!node.is_leaf) {
          stack[sp++] = node.b;
          stack[sp++] = node.a;
        } else {
          for (uint32_t k = node.a; k < node.b && n_obj < 64; ++k)
            objects[n_obj++] = k;
        }
      }
    }
    if (n_obj == 0)
      break;
    for (uint32_t k = 0; k < n_obj; ++k) {
      uint32_t orig = s.bvh.perm[objects[k]];
      if (c)
        ++c->tris;
      T t;
      uint32_t back;
      if (intersect_triangle(ray, s.tris[orig], &t, &back) &&
          (best.idx == kNoHit || hit_before(t, orig, best.t, best.idx)))
        best = {t, orig, back};
    }
    n_obj = 0;
  }
  return best;
}

template <class T> struct RenderArgs {
  const OScene<T> *scene;
  IntegratorSettings settings;
  WorkDivision wd;
  T film_to_world[12];
  const float *sobol; // vals[sample * dim_bound + dim]
  uint32_t dim_bound, initial_dim_bound, samples_per;
  uint32_t rng_type;
  // InitialRays / InitialIdxAndDir sample specs (initial_ray_sample.h:16-75):
  // one ray (o, d) per "pixel" x; with initial_idx the first intersection is
  // that triangle at distance 1, front-facing
  const float *initial_rays = nullptr;
  const uint32_t *initial_idx = nullptr;
  // OutputPerStep (rendering_equation.h:56-123): radiance added by iteration
  // k goes to step image k; paths stop after n_steps iterations
  uint32_t n_steps = 0;
  // First-hit G-buffer of sample 0 per pixel (12 floats: position, shading
  // normal, albedo summary, emission; zeros on a miss), as the batched
  // datagen renders write it.
  T *gbuffer = nullptr;
};

template <class T> struct TableRngT {
  const float *vals;
  uint32_t row, dim, dim_bound;
  T next() {
    T out = T(vals[row + dim]);
    dim = dim + 1 == dim_bound ? 0 : dim + 1;
    return out;
  }
};

template <class T> struct HashRngT {
  HashRng r;
  T next() { return T(r.next()); }
};

template <class T, class R>
V3<T> integrate_thread(const RenderArgs<T> &a, const ThreadInfo &info,
                       uint32_t location, KindCounts *cnt, V3<T> *steps) {
  V3<T> total = {T(0), T(0), T(0)};
  const SceneView<T> view = a.scene->view();
  const bool per_step = a.n_steps != 0;
  for (uint32_t sample = info.start_sample; sample < info.end_sample; ++sample) {
    R rng;
    if constexpr (std::is_same_v<R, TableRngT<T>>)
      rng = {a.sobol, sample * a.dim_bound,
             initial_dimension(location, a.initial_dim_bound), a.dim_bound};
    else
      rng = {make_hash_rng(sample + location * a.samples_per)};
    Ray<T> next_ray;
    Hit<T> given{T(0), kNoHit, 0};
    if (a.initial_rays) {
      const float *r = a.initial_rays + 6 * size_t(info.x);
      next_ray = {{T(r[0]), T(r[1]), T(r[2])}, {T(r[3]), T(r[4]), T(r[5])}};
      if (a.initial_idx)
        given = {T(1), a.initial_idx[info.x], 0};
    } else {
      T xo = rng.next();
      T yo = rng.next();
      next_ray = camera_ray(a.film_to_world, T(float(info.x)) + xo, T(float(info.y)) + yo,
                            a.wd.x_dim, a.wd.y_dim);
    }
    Ray<T> light_ray = next_ray;
    PathState<T> st = initial_path_state<T>();
    for (uint32_t step = 0;; ++step) {
      if (per_step) {
        if (step >= a.n_steps)
          break;
        st.total = {T(0), T(0), T(0)};
      }
      Hit<T> light_hit{T(0), kNoHit, 0}, next_hit{T(0), kNoHit, 0};
      if (st.n_light) {
        light_hit = bvh_closest(*a.scene, light_ray, &cnt->c[1]);
        cnt->rays[1] += 1;
      }
      if (st.has_next) {
        if (step == 0 && a.initial_idx) {
          next_hit = given;
        } else {
          next_hit = bvh_closest(*a.scene, next_ray, &cnt->c[0]);
          cnt->rays[0] += 1;
        }
      }
      PathState<T> ns;
      Ray<T> nl = light_ray, nn = next_ray;
      VertexInfo<T> vi{};
      const bool want_g = a.gbuffer && sample == 0 && step == 0;
      const bool done = rendering_equation_iteration(st, rng, next_ray, light_hit, next_hit,
                                                     a.settings, view, &ns, &nl, &nn,
                                                     want_g ? &vi : nullptr);
      if (want_g && vi.valid) {
        const Material<T> &m = view.materials[vi.material];
        const V3<T> alb = material_albedo(m);
        T *g = a.gbuffer + 12 * size_t(location);
        const V3<T> vals[4] = {vi.p, vi.normal, alb, m.emission};
        for (int k = 0; k < 4; ++k)
          g[3 * k] = vals[k].x, g[3 * k + 1] = vals[k].y, g[3 * k + 2] = vals[k].z;
      }
      if (per_step)
        steps[step] += ns.total;
      if (done) {
        if (!per_step)
          total += ns.total;
        break;
      }
      st = ns;
      light_ray = nl;
      next_ray = nn;
    }
  }
  return total;
}

// ---------------------------------------------------------------------------
// FP64 forward-mode derivative of the mega-kernel estimator for fixed samples
// (the check of the GPU's reverse-mode path replay). The path loop of
// rendering_equation.h:27-129 / rendering_equation_iteration.h:9-153 is
// restated here with, next to every multiplier, its derivative with respect
// to each material's albedo color (one channel at a time: channel c of a
// multiplier depends only on channel c of the albedos). The albedo-linear
// factors are evaluated twice, with the material's color and with a unit
// color, so no quantity is ever divided by an albedo:
//   vertex of material m:  W = sample weight,  Wu = weight with unit color
//     d(M * W)[j] = dM[j] * W + [j == m] * M * Wu     (same for the NEE eval)
//   emission term  M * Le[e]:  dL/dalbedo[j] += dM[j] * Le[e],
//                              dL/dLe[e]     += M.
// The emitter-selection CDF is held fixed (as it is exact when one material
// emits).
// ---------------------------------------------------------------------------
constexpr uint32_t kOracleMaxMats = 64;

using V3d = V3<double>;

Material<double> unit_color(Material<double> m) {
  if (m.type == kDiffuse)
    m.diffuse = {1.0, 1.0, 1.0};
  else
    m.specular = {1.0, 1.0, 1.0};
  return m;
}

// Returns false if the unit-color replay of a direction sample diverged
// (the sampler is expected not to depend on the color).
template <class R>
bool adjoint_path(const RenderArgs<double> &a, uint32_t px, uint32_t py, R rng,
                  const double d_i[3], uint32_t n_mats, double *g_alb, double *g_em,
                  V3d *radiance) {
  const SceneView<double> view = a.scene->view();
  const IntegratorSettings &s = a.settings;
  double xo = rng.next();
  double yo = rng.next();
  Ray<double> ray = camera_ray(a.film_to_world, double(float(px)) + xo,
                               double(float(py)) + yo, a.wd.x_dim, a.wd.y_dim);
  Ray<double> light_ray = ray;
  V3d M = {1.0, 1.0, 1.0}, LM = {0.0, 0.0, 0.0}, total = {0.0, 0.0, 0.0};
  V3d dM[kOracleMaxMats], dLM[kOracleMaxMats];
  for (uint32_t j = 0; j < n_mats; ++j)
    dM[j] = dLM[j] = {0.0, 0.0, 0.0};
  double light_target = 0.0;
  bool has_light = false, has_next = true, count_emission = true;
  const bool performs_samples = s.light_sampler != kLightNone;
  auto add_emission = [&](const V3d &mult, const V3d *dmult, uint32_t e) {
    const V3d le = view.materials[e].emission;
    total += mult * le;
    for (uint32_t j = 0; j < n_mats; ++j) {
      const V3d d = dmult[j] * le;
      g_alb[3 * j] += d_i[0] * d.x;
      g_alb[3 * j + 1] += d_i[1] * d.y;
      g_alb[3 * j + 2] += d_i[2] * d.z;
    }
    g_em[3 * e] += d_i[0] * mult.x;
    g_em[3 * e + 1] += d_i[1] * mult.y;
    g_em[3 * e + 2] += d_i[2] * mult.z;
  };
  for (uint32_t iters = 0;; ++iters) {
    Hit<double> light_hit{0.0, kNoHit, 0}, next_hit{0.0, kNoHit, 0};
    if (has_light)
      light_hit = bvh_closest(*a.scene, light_ray, nullptr);
    if (has_next)
      next_hit = bvh_closest(*a.scene, ray, nullptr);
    if (has_light && light_hit.idx != kNoHit && within_target(light_target, light_hit.t) &&
        !(s.back_cull_emission && light_hit.back))
      add_emission(LM, dLM, view.tris[light_hit.idx].material);
    if (!has_next || next_hit.idx == kNoHit)
      break;
    const V3d p = hit_point(ray, next_hit.t);
    const ShadeTri<double> tri = view.tris[next_hit.idx];
    const uint32_t m = tri.material;
    const Material<double> mat = view.materials[m];
    const Material<double> unit = unit_color(mat);
    if ((!performs_samples || count_emission) && !(s.back_cull_emission && next_hit.back))
      add_emission(M, dM, m);
    const TriGeom<double> g = {tri.v0, tri.e1, tri.e2};
    const V3d vn[3] = {tri.n0, tri.n1, tri.n2};
    const V3d normal = shading_normal(g, vn, p);
    const V3d in = ray.d;
    has_light = false;
    if (performs_samples) {
      const double pc = prob_continuous(mat);
      LightSample<double> ls;
      if (pc > 0.0 && sample_light(view, p, mat, normal, rng, &ls)) {
        const V3d e = continuous_eval(mat, in, ls.dir, normal);
        const V3d eu = continuous_eval(unit, in, ls.dir, normal);
        LM = ((M * e) * pc) / ls.multiplier;
        for (uint32_t j = 0; j < n_mats; ++j)
          dLM[j] = ((dM[j] * e) * pc) / ls.multiplier;
        dLM[m] += ((M * eu) * pc) / ls.multiplier;
        light_target = ls.target_distance;
        has_light = true;
        light_ray = {p, ls.dir};
      }
    }
    R rng_unit = rng;
    const DirSample<double> ds = sample_direction(s, mat, in, normal, rng);
    const DirSample<double> du = sample_direction(s, unit, in, normal, rng_unit);
    if (!(ds.dir.x == du.dir.x && ds.dir.y == du.dir.y && ds.dir.z == du.dir.z))
      return false;
    count_emission = ds.is_discrete;
    V3d nm = M * ds.multiplier;
    V3d dnm[kOracleMaxMats];
    for (uint32_t j = 0; j < n_mats; ++j)
      dnm[j] = dM[j] * ds.multiplier;
    dnm[m] += M * du.multiplier;
    const double tp = term_prob(s, iters, nm);
    has_next = rng.next() > tp;
    if (has_next) {
      const double denom = 1.0 - tp;
      M = {nm.x / denom, nm.y / denom, nm.z / denom};
      for (uint32_t j = 0; j < n_mats; ++j)
        dM[j] = {dnm[j].x / denom, dnm[j].y / denom, dnm[j].z / denom};
      ray = {p, ds.dir};
    } else if (!has_light) {
      break;
    }
  }
  *radiance = total;
  return true;
}

// Renders the pixels [pixel_begin, pixel_end) (row-major) or all when
// pixel_end == 0. Output: per pixel RGB (already divided by spp when there is a
// single sample block, else the sum over sample blocks divided by spp).
template <class T>
int render_impl(const RenderArgs<T> &a, uint32_t pixel_begin, uint32_t pixel_end,
                int n_threads, T *out, uint64_t *counts /* 6 */) {
  const WorkDivision &wd = a.wd;
  const uint32_t n_pixels = wd.x_dim * wd.y_dim;
  if (pixel_end == 0 || pixel_end > n_pixels)
    pixel_end = n_pixels;
  const uint32_t n_images = a.n_steps ? a.n_steps : 1;
  const size_t image = size_t(n_pixels) * wd.num_sample_blocks * 3;
  std::vector<T> partial(image * n_images, T(0));
  std::atomic<uint64_t> acc[6];
  for (auto &x : acc)
    x = 0;
  const int64_t total_blocks = wd.total_num_blocks();
  if (n_threads > 0)
    omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t block = 0; block < total_blocks; ++block) {
    // [image][thread]
    std::vector<V3<T>> thread_tot(size_t(wd.block_size) * n_images,
                                  V3<T>{T(0), T(0), T(0)});
    std::vector<V3<T>> steps(n_images);
    bool any = false;
    KindCounts kc;
    for (uint32_t th = 0; th < wd.block_size; ++th) {
      ThreadInfo info = wd.get_thread_info(uint32_t(block), th);
      if (info.exit)
        continue;
      uint32_t loc = info.x + info.y * wd.x_dim;
      if (loc < pixel_begin || loc >= pixel_end)
        continue;
      any = true;
      for (auto &v : steps)
        v = {T(0), T(0), T(0)};
      V3<T> t;
      if (a.rng_type == kRngSobel)
        t = integrate_thread<T, TableRngT<T>>(a, info, loc, &kc, steps.data());
      else
        t = integrate_thread<T, HashRngT<T>>(a, info, loc, &kc, steps.data());
      if (a.n_steps)
        for (uint32_t i = 0; i < n_images; ++i)
          thread_tot[size_t(i) * wd.block_size + th] = steps[i];
      else
        thread_tot[th] = t;
    }
    if (!any)
      continue;
    uint32_t sbs = wd.sample_block_size;
    for (uint32_t i = 0; i < n_images; ++i)
      for (uint32_t th = 0; th < wd.block_size; th += sbs) {
        ThreadInfo info = wd.get_thread_info(uint32_t(block), th);
        if (info.exit)
          continue;
        V3<T> sum = tree_reduce(thread_tot.data() + size_t(i) * wd.block_size + th, sbs);
        size_t idx = size_t(wd.sample_block_idx(uint32_t(block))) +
                     size_t(wd.num_sample_blocks) * (info.x + size_t(info.y) * wd.x_dim);
        T *pp = partial.data() + i * image;
        pp[3 * idx] = sum.x;
        pp[3 * idx + 1] = sum.y;
        pp[3 * idx + 2] = sum.z;
      }
    for (int k = 0; k < 2; ++k) {
      acc[3 * k] += kc.rays[k];
      acc[3 * k + 1] += kc.c[k].nodes;
      acc[3 * k + 2] += kc.c[k].tris;
    }
  }
  T spp = T(float(a.samples_per));
  for (uint32_t i = 0; i < n_images; ++i) {
    const T *pp = partial.data() + i * image;
    T *o = out + size_t(i) * n_pixels * 3;
    for (uint32_t p = pixel_begin; p < pixel_end; ++p) {
      T s[3] = {T(0), T(0), T(0)};
      for (uint32_t k = 0; k < wd.num_sample_blocks; ++k)
        for (int ch = 0; ch < 3; ++ch)
          s[ch] = wd.num_sample_blocks == 1
                      ? pp[3 * (size_t(p) * wd.num_sample_blocks + k) + ch]
                      : s[ch] + pp[3 * (size_t(p) * wd.num_sample_blocks + k) + ch];
      for (int ch = 0; ch < 3; ++ch)
        o[3 * size_t(p) + ch] = s[ch] / spp;
    }
  }
  if (counts)
    for (int k = 0; k < 6; ++k)
      counts[k] = acc[k];
  return 0;
}

template <class T>
void fill_scene(OScene<T> &s, const float *tris, const float *normals,
                const uint32_t *tri_mat, uint32_t n_tris, const pt_material *mats,
                uint32_t n_mats, const float *light_tris, const float *light_cdf,
                uint32_t n_lights) {
  s.tris.resize(n_tris);
  s.shade.resize(n_tris);
  for (uint32_t i = 0; i < n_tris; ++i) {
    s.tris[i] = load_tri<T>(tris + 9 * i);
    V3<T> n[3];
    for (int k = 0; k < 3; ++k)
      n[k] = {T(normals[9 * i + 3 * k]), T(normals[9 * i + 3 * k + 1]),
              T(normals[9 * i + 3 * k + 2])};
    s.shade[i] = make_shade_tri(s.tris[i], n, tri_mat[i]);
  }
  for (uint32_t i = 0; i < n_mats; ++i) {
    const pt_material &m = mats[i];
    s.mats.push_back(make_material<T>(
        m.type, V3<T>{T(m.diffuse[0]), T(m.diffuse[1]), T(m.diffuse[2])},
        V3<T>{T(m.specular[0]), T(m.specular[1]), T(m.specular[2])},
        T(m.shininess), T(m.ior), T(m.diffuse_weight), T(m.glossy_weight),
        V3<T>{T(m.emission[0]), T(m.emission[1]), T(m.emission[2])}));
  }
  for (uint32_t i = 0; i < n_lights; ++i) {
    s.light_tris.push_back(load_tri<T>(light_tris + 9 * i));
    s.light_cdf.push_back(T(light_cdf[i]));
  }
  build_accel(s, float_tris(tris, n_tris));
}

IntegratorSettings integrator_from(const pt_settings &st) {
  IntegratorSettings is{};
  is.light_sampler = st.light_sampler;
  is.dir_sampler = st.dir_sampler;
  is.term_prob = st.term_prob;
  is.rng = st.rng;
  is.back_cull_emission = st.back_cull_emission;
  is.n_iters = st.n_iters;
  is.constant_prob = st.constant_prob;
  is.mf_exp = st.multiplier_func_exp;
  is.mf_min_prob = st.multiplier_func_min_prob;
  is.normalize_min_prob = st.normalize_min_prob;
  return is;
}

} // namespace

extern "C" {

// Brute-force closest hit (float). rays: n*6.
int oracle_loopall(const float *tris, uint32_t n_tris, const float *rays,
                   uint32_t n, float *t, uint32_t *idx, uint32_t *back) {
  std::vector<TriGeom<float>> tv = float_tris(tris, n_tris);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    const float *r = rays + 6 * i;
    Ray<float> ray = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
    Hit<float> h = loop_all(tv, ray);
    t[i] = h.t;
    idx[i] = h.idx;
    back[i] = h.back;
  }
  return 0;
}

// Reference BVH traversal over the host SAH tree (float), with per-ray counts.
int oracle_bvh_intersect(const float *tris, uint32_t n_tris, const float *rays,
                         uint32_t n, float *t, uint32_t *idx, uint32_t *back,
                         uint32_t *nodes_visited, uint32_t *tris_tested) {
  OScene<float> s;
  s.tris = float_tris(tris, n_tris);
  build_accel(s, s.tris);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    const float *r = rays + 6 * i;
    Ray<float> ray = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
    Counts c;
    Hit<float> h = bvh_closest(s, ray, &c);
    t[i] = h.t;
    idx[i] = h.idx;
    back[i] = h.back;
    if (nodes_visited)
      nodes_visited[i] = uint32_t(c.nodes);
    if (tris_tested)
      tris_tested[i] = uint32_t(c.tris);
  }
  return 0;
}

// Sobol table, direct evaluation of each point. dirs: dim_bound*32.
int oracle_sobol_table(const uint32_t *dirs, uint32_t dim_bound, uint32_t count,
                       float *out) {
  for (uint32_t s = 0; s < count; ++s)
    for (uint32_t d = 0; d < dim_bound; ++d)
      out[size_t(s) * dim_bound + d] = sobol_table_value(dirs, d, s);
  return 0;
}

uint32_t oracle_fnv_hash(uint32_t v) { return fnv_hash(v); }

int oracle_work_division(const pt_work_division_settings *s, uint32_t spp,
                         uint32_t x, uint32_t y, uint32_t *thread_info_out,
                         uint32_t max_threads) {
  // Writes (x, y, start, end, exit) per (block, thread) in launch order.
  WorkDivisionSettings ws;
  ws.block_size = s->block_size;
  ws.target_x_block_size = s->target_x_block_size;
  ws.force_target_samples = s->force_target_samples;
  ws.forced_target_samples_per_thread = s->forced_target_samples_per_thread;
  ws.base_num_threads = s->base_num_threads;
  ws.samples_per_thread_scaling_power = s->samples_per_thread_scaling_power;
  ws.max_samples_per_thread = s->max_samples_per_thread;
  WorkDivision wd;
  if (!WorkDivision::make(ws, spp, x, y, &wd))
    return -1;
  uint64_t total = uint64_t(wd.total_num_blocks()) * wd.block_size;
  if (total > max_threads)
    return int(total > 0x7fffffff ? 0x7fffffff : total);
  for (uint32_t b = 0; b < wd.total_num_blocks(); ++b)
    for (uint32_t t = 0; t < wd.block_size; ++t) {
      ThreadInfo i = wd.get_thread_info(b, t);
      uint32_t *o = thread_info_out + 5 * (size_t(b) * wd.block_size + t);
      o[0] = i.x, o[1] = i.y, o[2] = i.start_sample, o[3] = i.end_sample,
      o[4] = i.exit;
    }
  return int(total);
}

// FP64 forward-mode gradients of sum(d_image * image) for the mega-kernel
// estimator with the same samples: grad_albedo / grad_emission per material
// (n_mats * 3 doubles each; albedo = diffuse for Diffuse materials, specular
// otherwise) and the FP64 image (x_dim * y_dim * 3). Returns 0, -1 on bad
// settings, -2 if a direction sampler depended on a color.
int oracle_adjoint(const float *tris, const float *normals, const uint32_t *tri_mat,
                   uint32_t n_tris, const pt_material *mats, uint32_t n_mats,
                   const float *light_tris, const float *light_cdf, uint32_t n_lights,
                   const float *film_to_world, const pt_settings *settings,
                   uint32_t samples_per, uint32_t x_dim, uint32_t y_dim,
                   const float *sobol, const double *d_image, int n_threads,
                   double *grad_albedo, double *grad_emission, double *image,
                   double *seconds) {
  if (n_mats == 0 || n_mats > kOracleMaxMats)
    return -1;
  WorkDivision wd{};
  wd.x_dim = x_dim;
  wd.y_dim = y_dim;
  IntegratorSettings is = integrator_from(*settings);
  uint32_t dim_bound = settings->max_sample_size;
  uint32_t init_bound = dim_bound < 64 ? dim_bound : 64;
  OScene<double> sc;
  fill_scene(sc, tris, normals, tri_mat, n_tris, mats, n_mats, light_tris, light_cdf,
             n_lights);
  RenderArgs<double> a{&sc, is, wd, {}, sobol, dim_bound, init_bound, samples_per,
                       settings->rng};
  for (int i = 0; i < 12; ++i)
    a.film_to_world[i] = film_to_world[i];
  if (n_threads > 0)
    omp_set_num_threads(n_threads);
  const uint32_t n_pixels = x_dim * y_dim;
  const int max_t = omp_get_max_threads();
  std::vector<double> ga(size_t(max_t) * n_mats * 3, 0.0), ge(size_t(max_t) * n_mats * 3, 0.0);
  std::atomic<int> bad{0};
  double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t pix = 0; pix < int64_t(n_pixels); ++pix) {
    const int th = omp_get_thread_num();
    double *g_alb = ga.data() + size_t(th) * n_mats * 3;
    double *g_em = ge.data() + size_t(th) * n_mats * 3;
    const uint32_t px = uint32_t(pix) % x_dim, py = uint32_t(pix) / x_dim;
    const uint32_t location = uint32_t(pix);
    const double d_i[3] = {d_image[3 * pix] / samples_per, d_image[3 * pix + 1] / samples_per,
                           d_image[3 * pix + 2] / samples_per};
    V3d sum = {0.0, 0.0, 0.0};
    for (uint32_t sample = 0; sample < samples_per; ++sample) {
      V3d rad;
      bool ok;
      if (settings->rng == kRngSobel)
        ok = adjoint_path(a, px, py,
                          TableRngT<double>{sobol, sample * dim_bound,
                                            initial_dimension(location, init_bound), dim_bound},
                          d_i, n_mats, g_alb, g_em, &rad);
      else
        ok = adjoint_path(a, px, py,
                          HashRngT<double>{make_hash_rng(sample + location * samples_per)},
                          d_i, n_mats, g_alb, g_em, &rad);
      if (!ok)
        bad = 1;
      sum += rad;
    }
    if (image) {
      image[3 * pix] = sum.x / samples_per;
      image[3 * pix + 1] = sum.y / samples_per;
      image[3 * pix + 2] = sum.z / samples_per;
    }
  }
  double t1 = omp_get_wtime();
  for (uint32_t i = 0; i < n_mats * 3; ++i) {
    double x = 0.0, y = 0.0;
    for (int t = 0; t < max_t; ++t) {
      x += ga[size_t(t) * n_mats * 3 + i];
      y += ge[size_t(t) * n_mats * 3 + i];
    }
    grad_albedo[i] = x;
    grad_emission[i] = y;
  }
  if (seconds)
    *seconds = t1 - t0;
  return bad ? -2 : 0;
}

// Mega-kernel-equivalent render. precision: 32 or 64. Output out32 (float) or
// out64 (double), x_dim*y_dim*3. Counts per render: rays traced, BVH nodes
// visited and triangles tested by the reference traversal.
int oracle_render_ex(const float *initial_rays, const uint32_t *initial_idx,
                     uint32_t n_steps, float *gbuffer, const float *tris, const float *normals, const uint32_t *tri_mat,
                  uint32_t n_tris, const pt_material *mats, uint32_t n_mats,
                  const float *light_tris, const float *light_cdf,
                  uint32_t n_lights, const float *film_to_world,
                  const pt_settings *settings, uint32_t samples_per,
                  uint32_t x_dim, uint32_t y_dim, const float *sobol,
                  int precision, int n_threads, uint32_t pixel_begin,
                  uint32_t pixel_end, float *out32, double *out64,
                  uint64_t *counts /* 6: per kind (camera/continuation,
                                      light): rays, nodes, tris */,
                  double *seconds /* render time, scene setup excluded */) {
  WorkDivisionSettings ws;
  const pt_work_division_settings &w = settings->render_work_division;
  ws.block_size = w.block_size;
  ws.target_x_block_size = w.target_x_block_size;
  ws.force_target_samples = w.force_target_samples;
  ws.forced_target_samples_per_thread = w.forced_target_samples_per_thread;
  ws.base_num_threads = w.base_num_threads;
  ws.samples_per_thread_scaling_power = w.samples_per_thread_scaling_power;
  ws.max_samples_per_thread = w.max_samples_per_thread;
  WorkDivision wd;
  if (!WorkDivision::make(ws, samples_per, x_dim, y_dim, &wd))
    return -1;
  IntegratorSettings is = integrator_from(*settings);
  uint32_t dim_bound = settings->max_sample_size;
  uint32_t init_bound = dim_bound < 64 ? dim_bound : 64;
  uint64_t c[6] = {0, 0, 0, 0, 0, 0};
  int rc;
  double t0 = 0, t1 = 0;
  if (precision == 64) {
    OScene<double> s;
    fill_scene(s, tris, normals, tri_mat, n_tris, mats, n_mats, light_tris,
               light_cdf, n_lights);
    RenderArgs<double> a{&s, is, wd, {}, sobol, dim_bound, init_bound,
                         samples_per, settings->rng, initial_rays, initial_idx, n_steps};
    for (int i = 0; i < 12; ++i)
      a.film_to_world[i] = film_to_world[i];
    t0 = omp_get_wtime();
    rc = render_impl(a, pixel_begin, pixel_end, n_threads, out64, c);
    t1 = omp_get_wtime();
  } else {
    OScene<float> s;
    fill_scene(s, tris, normals, tri_mat, n_tris, mats, n_mats, light_tris,
               light_cdf, n_lights);
    RenderArgs<float> a{&s, is, wd, {}, sobol, dim_bound, init_bound,
                        samples_per, settings->rng, initial_rays, initial_idx, n_steps,
                        gbuffer};
    for (int i = 0; i < 12; ++i)
      a.film_to_world[i] = film_to_world[i];
    t0 = omp_get_wtime();
    rc = render_impl(a, pixel_begin, pixel_end, n_threads, out32, c);
    t1 = omp_get_wtime();
  }
  if (counts)
    std::memcpy(counts, c, sizeof(c));
  if (seconds)
    *seconds = t1 - t0;
  return rc;
}

int oracle_render(const float *tris, const float *normals, const uint32_t *tri_mat,
                  uint32_t n_tris, const pt_material *mats, uint32_t n_mats,
                  const float *light_tris, const float *light_cdf,
                  uint32_t n_lights, const float *film_to_world,
                  const pt_settings *settings, uint32_t samples_per,
                  uint32_t x_dim, uint32_t y_dim, const float *sobol,
                  int precision, int n_threads, uint32_t pixel_begin,
                  uint32_t pixel_end, float *out32, double *out64,
                  uint64_t *counts /* 6: per kind (camera/continuation,
                                      light): rays, nodes, tris */,
                  double *seconds /* render time, scene setup excluded */) {
  return oracle_render_ex(nullptr, nullptr, 0, nullptr, tris, normals, tri_mat, n_tris, mats, n_mats,
                          light_tris, light_cdf, n_lights, film_to_world, settings,
                          samples_per, x_dim, y_dim, sobol, precision, n_threads,
                          pixel_begin, pixel_end, out32, out64, counts, seconds);
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

namespace {
WorkDivisionSettings wd_settings(const pt_work_division_settings &w) {
  WorkDivisionSettings ws;
  ws.block_size = w.block_size;
  ws.target_x_block_size = w.target_x_block_size;
  ws.force_target_samples = w.force_target_samples;
  ws.forced_target_samples_per_thread = w.forced_target_samples_per_thread;
  ws.base_num_threads = w.base_num_threads;
  ws.samples_per_thread_scaling_power = w.samples_per_thread_scaling_power;
  ws.max_samples_per_thread = w.max_samples_per_thread;
  return ws;
}
} // namespace

// The reduction of the reference's reduce test (tests/src/reduce.cu:49-158):
// every work-division thread sums in[sample + location * samples_per] over
// its sample range in order, then the threads of a location are combined by
// the lane reduction tree. out[location] (num_locations = x_dim, y_dim = 1).
// Returns -1 on invalid settings or more than one sample block.
int oracle_reduce_samples(const pt_work_division_settings *s, const float *in,
                          uint32_t num_locations, uint32_t samples_per, float *out) {
  WorkDivision wd;
  if (!WorkDivision::make(wd_settings(*s), samples_per, num_locations, 1, &wd) ||
      wd.num_sample_blocks != 1)
    return -1;
  std::vector<float> vals(wd.block_size);
  for (uint32_t b = 0; b < wd.total_num_blocks(); ++b) {
    for (uint32_t t = 0; t < wd.block_size; ++t) {
      ThreadInfo info = wd.get_thread_info(b, t);
      float total = 0.f;
      if (!info.exit)
        for (uint32_t i = info.start_sample; i < info.end_sample; ++i)
          total += in[i + size_t(info.x) * samples_per];
      vals[t] = total;
    }
    for (uint32_t t = 0; t < wd.block_size; t += wd.sample_block_size) {
      ThreadInfo info = wd.get_thread_info(b, t);
      if (!info.exit)
        out[info.x] = tree_reduce(vals.data() + t, wd.sample_block_size);
    }
  }
  return 0;
}

// The coverage walk of the reference's work-division test
// (tests/src/work_division.cpp:24-80): threads in launch order must advance
// through samples, then x within an x block, then y within a y block, then
// the next block, with exit set exactly outside the image, and the dimension
// check of :82-91. Returns 0, or the 1-based index of the first failing
// check * 2^32 + thread ordinal (0 for the end-of-walk checks).
uint64_t oracle_work_division_walk(const pt_work_division_settings *s, uint32_t samples_per,
                                   uint32_t x_dim, uint32_t y_dim) {
  WorkDivision d;
  if (!WorkDivision::make(wd_settings(*s), samples_per, x_dim, y_dim, &d))
    return uint64_t(100) << 32;
  const uint32_t spp = d.num_sample_blocks * d.sample_block_size * d.base_samples_per_thread +
                       d.n_threads_per_unit_extra;
  const uint32_t ex = d.x_block_size * d.num_x_blocks, ey = d.y_block_size * d.num_y_blocks;
  if (spp != samples_per)
    return uint64_t(101) << 32;
  if (!(ex >= x_dim && ey >= y_dim && ex < x_dim + d.x_block_size &&
        ey < y_dim + d.y_block_size))
    return uint64_t(102) << 32;
  uint32_t samples_up_to = 0, x_up_to = 0, y_up_to = 0;
  uint64_t ordinal = 0;
  auto fail = [&](uint64_t code) { return (code << 32) | ordinal; };
  for (uint32_t b = 0; b < d.total_num_blocks(); ++b) {
    for (uint32_t t = 0; t < d.block_size; ++t, ++ordinal) {
      const ThreadInfo info = d.get_thread_info(b, t);
      if (bool(info.exit) != (info.x >= x_dim || info.y >= y_dim))
        return fail(1);
      if (info.start_sample != samples_up_to)
        return fail(2);
      if (info.x != x_up_to)
        return fail(3);
      if (info.y != y_up_to)
        return fail(4);
      samples_up_to = info.end_sample;
      if (samples_up_to > spp)
        return fail(5);
      if (samples_up_to == spp && (t + 1) % d.sample_block_size == 0) {
        ++x_up_to;
        samples_up_to = 0;
        if (x_up_to % d.x_block_size == 0) {
          if (t == d.block_size - 1) {
            if (x_up_to == ex) {
              if ((b + 1) % d.num_x_blocks != 0)
                return fail(6);
              ++y_up_to;
              x_up_to = 0;
            } else {
              if (y_up_to < d.y_block_size - 1)
                return fail(7);
              y_up_to -= d.y_block_size - 1;
            }
          } else {
            ++y_up_to;
            if (x_up_to < d.x_block_size)
              return fail(8);
            x_up_to -= d.x_block_size;
          }
        }
      }
      if (x_up_to > ex || y_up_to > ey)
        return fail(9);
    }
  }
  if (samples_up_to != 0 || x_up_to != 0 || y_up_to != ey)
    return uint64_t(10) << 32;
  return 0;
}

// FloatRGB -> BGRA32, restated from include/lib/bgra_32.h:12-31 and
// include/lib/tone_map.h:7-10: per channel v / (v + 1) in float, times 255.0f,
// converted to uint8 by truncation; bytes in memory order B, G, R, 0.
void oracle_to_bgra(const float *rgb, uint32_t n, uint8_t *out) {
  for (uint32_t i = 0; i < n; ++i) {
    uint8_t c[3];
    for (int k = 0; k < 3; ++k) {
      const float v = rgb[3 * size_t(i) + k];
      const float mapped = v / (v + 1.0f);
      const float scaled = mapped * 255.0f;
      c[k] = static_cast<uint8_t>(scaled);
    }
    out[4 * size_t(i) + 0] = c[2];
    out[4 * size_t(i) + 1] = c[1];
    out[4 * size_t(i) + 2] = c[0];
    out[4 * size_t(i) + 3] = 0;
  }
}

// Values of the Uniform-rng stand-in for one seed (sample + location * spp).
void oracle_hash_rng(uint32_t seed, uint32_t n, float *out) {
  HashRng r = make_hash_rng(seed);
  for (uint32_t i = 0; i < n; ++i)
    out[i] = r.next();
}

namespace {
struct ArrayRng {
  const double *v;
  uint32_t n, used;
  double next() { return used < n ? v[used++] : 0.0; }
};
struct ArrayRngF {
  const float *v;
  uint32_t n, used;
  float next() { return used < n ? v[used++] : 0.0f; }
};
Material<double> to_material(const pt_material &m) {
  return make_material<double>(
      m.type, V3<double>{m.diffuse[0], m.diffuse[1], m.diffuse[2]},
      V3<double>{m.specular[0], m.specular[1], m.specular[2]}, m.shininess,
      m.ior, m.diffuse_weight, m.glossy_weight,
      V3<double>{m.emission[0], m.emission[1], m.emission[2]});
}
} // namespace

// One BSDF direction sample in double precision, driven by the given uniform
// values. out: dir[3], multiplier[3], is_discrete, values consumed.
void oracle_bsdf_sample(const pt_material *mat, const double *in, const double *n,
                        const double *u, uint32_t n_u, int dir_sampler, double *out) {
  IntegratorSettings s{};
  s.dir_sampler = uint32_t(dir_sampler);
  ArrayRng r{u, n_u, 0};
  DirSample<double> d = sample_direction(s, to_material(*mat), V3<double>{in[0], in[1], in[2]},
                                         V3<double>{n[0], n[1], n[2]}, r);
  out[0] = d.dir.x, out[1] = d.dir.y, out[2] = d.dir.z;
  out[3] = d.multiplier.x, out[4] = d.multiplier.y, out[5] = d.multiplier.z;
  out[6] = d.is_discrete ? 1.0 : 0.0;
  out[7] = double(r.used);
}

// The same in float (the GPU's precision): inputs and draws rounded to float.
void oracle_bsdf_sample_f32(const pt_material *mat, const double *in, const double *n,
                            const double *u, uint32_t n_u, int dir_sampler, double *out) {
  IntegratorSettings s{};
  s.dir_sampler = uint32_t(dir_sampler);
  float uf[16];
  const uint32_t m = n_u < 16 ? n_u : 16;
  for (uint32_t i = 0; i < m; ++i)
    uf[i] = float(u[i]);
  ArrayRngF r{uf, m, 0};
  const pt_material &pm = *mat;
  const Material<float> mf = make_material<float>(
      pm.type, V3<float>{pm.diffuse[0], pm.diffuse[1], pm.diffuse[2]},
      V3<float>{pm.specular[0], pm.specular[1], pm.specular[2]}, pm.shininess, pm.ior,
      pm.diffuse_weight, pm.glossy_weight,
      V3<float>{pm.emission[0], pm.emission[1], pm.emission[2]});
  DirSample<float> d = sample_direction(
      s, mf, V3<float>{float(in[0]), float(in[1]), float(in[2])},
      V3<float>{float(n[0]), float(n[1]), float(n[2])}, r);
  out[0] = d.dir.x, out[1] = d.dir.y, out[2] = d.dir.z;
  out[3] = d.multiplier.x, out[4] = d.multiplier.y, out[5] = d.multiplier.z;
  out[6] = d.is_discrete ? 1.0 : 0.0;
  out[7] = double(r.used);
}

// BSDF evaluation f(in, out) in double precision (continuous part) and the
// probability of the continuous lobe.
void oracle_bsdf_eval(const pt_material *mat, const double *in, const double *o,
                      const double *n, double *out) {
  Material<double> m = to_material(*mat);
  V3<double> e = continuous_eval(m, V3<double>{in[0], in[1], in[2]},
                                 V3<double>{o[0], o[1], o[2]}, V3<double>{n[0], n[1], n[2]});
  out[0] = e.x, out[1] = e.y, out[2] = e.z;
  out[3] = prob_continuous(m);
}

// Termination probability (double).
double oracle_term_prob(const pt_settings *st, uint32_t iters, const double *mult) {
  IntegratorSettings s{};
  s.term_prob = st->term_prob;
  s.n_iters = st->n_iters;
  s.constant_prob = st->constant_prob;
  s.mf_exp = st->multiplier_func_exp;
  s.mf_min_prob = st->multiplier_func_min_prob;
  s.normalize_min_prob = st->normalize_min_prob;
  return term_prob(s, iters, V3<double>{mult[0], mult[1], mult[2]});
}

// Deterministic math exposed for accuracy tests against libm.
void oracle_detmath(int fn, const float *x, const float *y, float *out,
                    uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    switch (fn) {
    case 0: {
      float s, c;
      detmath::sincos_2pi(x[i], &s, &c);
      out[2 * i] = s;
      out[2 * i + 1] = c;
      break;
    }
    case 1:
      out[i] = detmath::log2(x[i]);
      break;
    case 2:
      out[i] = detmath::exp2(x[i]);
      break;
    default:
      out[i] = detmath::pow(x[i], y[i]);
    }
  }
}

} // extern "C"
