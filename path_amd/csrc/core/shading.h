// BSDFs, light sampling, termination probabilities and one step of the
// rendering equation, shared by the GPU integrators and the CPU checker.
//
// Trigonometric round trips of the reference (asin/acos followed by sin/cos
// of the same angle, and the AngleAxis rotation built from acos of a dot
// product) are replaced by the identical (cos, sin) pairs computed directly,
// e.g. cos(asin(sqrt(v))) = sqrt(1 - v); the estimators are unchanged.
#pragma once

#include "detmath.h"
#include "geometry.h"

namespace pt {

// ---------------------------------------------------------------------------
// Settings (reference: include/render/settings.h:27-45 and the per-plugin
// settings structs). Plain POD so it can be passed to kernels by value.
// ---------------------------------------------------------------------------
enum LightSamplerType : uint32_t { kLightNone = 0, kLightRandomTriangle = 1 };
enum DirSamplerType : uint32_t { kDirUniform = 0, kDirBSDF = 1 };
enum TermProbType : uint32_t {
  kTermConstant = 0,
  kTermNIters = 1,
  kTermMultiplierFunc = 2,
  kTermNormalize = 3,
};
enum RngType : uint32_t { kRngUniform = 0, kRngSobel = 1 };

struct IntegratorSettings {
  uint32_t light_sampler;      // LightSamplerType
  uint32_t dir_sampler;        // DirSamplerType
  uint32_t term_prob;          // TermProbType
  uint32_t rng;                // RngType
  uint32_t back_cull_emission; // reference default true
  uint32_t n_iters;            // NIters::iters (reference default 1)
  float constant_prob;         // Constant::prob
  float mf_exp;                // MultiplierFunc::exp (default 50)
  float mf_min_prob;           // MultiplierFunc::min_prob (default 0.1)
  float normalize_min_prob;    // Normalize::min_prob
};

// ---------------------------------------------------------------------------
// Materials: reference include/bsdf/union_bsdf.h:17-88 (tagged union) and
// include/scene/tinyobj_material_conversion.h:13-101 (parameters).
// ---------------------------------------------------------------------------
enum BSDFType : uint32_t {
  kDiffuse = 0,
  kGlossy = 1,
  kDiffuseGlossy = 2,
  kDiffuseMirror = 3,
  kMirror = 4,
  kDielectricRefractive = 5,
};

template <class T> struct Material {
  uint32_t type;
  uint32_t pad_;
  V3<T> diffuse;  // Diffuse albedo (the reweighted one inside Combined)
  V3<T> specular; // Glossy / Mirror / Dielectric coefficient
  T shininess;
  T glossy_norm;      // (shininess + 2) / (2 pi)
  T inv_shininess_p1; // 1 / (shininess + 1)
  T ior;
  T r0;      // ((ior - 1) / (ior + 1))^2
  T weight0; // DiffuseGlossy: inclusive weight of the diffuse lobe;
             // DiffuseMirror: probability of the continuous (diffuse) lobe
  V3<T> emission;
};

// Build the derived constants exactly as the reference constructors do
// (include/bsdf/glossy.h:20-24, dielectric_refractive.h:18-22,
// combined.h:22-62).
template <class T>
PT_HD inline Material<T> make_material(uint32_t type, V3<T> diffuse,
                                       V3<T> specular, T shininess, T ior,
                                       T diffuse_weight, T glossy_weight,
                                       V3<T> emission) {
  Material<T> m{};
  m.type = type;
  m.diffuse = diffuse;
  m.specular = specular;
  m.shininess = shininess;
  m.glossy_norm = (shininess + T(2)) / (T(2) * T(3.14159265358979323846));
  m.inv_shininess_p1 = T(1) / (shininess + T(1));
  m.ior = ior;
  T r = (ior - T(1)) / (ior + T(1));
  m.r0 = r * r;
  if (type == kDiffuseGlossy) {
    m.weight0 = diffuse_weight / (diffuse_weight + glossy_weight);
  } else if (type == kDiffuseMirror) {
    m.weight0 = diffuse_weight;
  } else {
    m.weight0 = T(1);
  }
  m.emission = emission;
  return m;
}

template <class T> PT_INLINE bool is_brdf(const Material<T> &m) {
  return m.type != kDielectricRefractive;
}

// reference: include/bsdf/utils.h:10-13
template <class T> PT_INLINE V3<T> reflect_over_normal(V3<T> vec, V3<T> n) {
  return normalize(vec + n * (T(2) * -dot(vec, n)));
}

// Rotate the local frame (z = `axis_vec`) direction with polar cosine/sine
// (ct, st) and azimuth 2*pi*phi_turns into world space.
// reference: include/lib/projection.h:9-32 (rotation taking +z to vec, built
// as AngleAxis(acos(vec.z), normalize(vec x z)).inverse()).
template <class T>
PT_INLINE V3<T> find_relative_vec(V3<T> vec, T phi_turns, T ct, T st) {
  T sp, cp;
  Math<T>::sincos_2pi(phi_turns, &sp, &cp);
  V3<T> local = {st * cp, st * sp, ct};
  T c = clamp_(vec.z, T(-1), T(1));
  T s = -sqrt_(max_(T(0), T(1) - c * c)); // sin of the inverse rotation
  V3<T> k = normalize(V3<T>{vec.y, -vec.x, T(0)});
  V3<T> r = (local * c + cross(k, local) * s) + k * (dot(k, local) * (T(1) - c));
  return normalize(r);
}

// reference: include/bsdf/utils.h:15-36
template <class T>
PT_INLINE V3<T> refract_by_normal(T ior, V3<T> vec, V3<T> n) {
  T cos_to_normal = -dot(vec, n);
  bool exiting = cos_to_normal < T(0);
  T frac = exiting ? ior : T(1) / ior;
  T in_sqrt = T(1) - (frac * frac) * (T(1) - cos_to_normal * cos_to_normal);
  V3<T> eff_n = exiting ? -n : n;
  if (in_sqrt > T(0)) {
    T ct = sqrt_(in_sqrt);
    T st = sqrt_(T(1) - in_sqrt);
    V3<T> k = normalize(cross(vec, eff_n));
    V3<T> v = -eff_n;
    V3<T> r = (v * ct + cross(k, v) * st) + k * (dot(k, v) * (T(1) - ct));
    return normalize(r);
  }
  return reflect_over_normal(vec, n);
}

// Phong lobe about the mirror direction. reference: include/bsdf/glossy.h:26-33
template <class T>
PT_INLINE V3<T> glossy_eval(const Material<T> &m, V3<T> in, V3<T> out,
                            V3<T> n) {
  T c = max_(dot(reflect_over_normal(in, n), out), T(0));
  return (m.specular * m.glossy_norm) * Math<T>::pow(c, m.shininess);
}

template <class T> PT_INLINE V3<T> diffuse_eval(const Material<T> &m) {
  return m.diffuse * T(0.318309886183790671538); // 1 / M_PI
}

// reference: include/bsdf/union_bsdf.h:41-50, combined.h:63-80
template <class T>
PT_INLINE V3<T> continuous_eval(const Material<T> &m, V3<T> in, V3<T> out,
                                V3<T> n) {
  switch (m.type) {
  case kGlossy:
    return glossy_eval(m, in, out, n);
  case kDiffuseGlossy: {
    V3<T> d = diffuse_eval(m) * m.weight0;
    V3<T> g = glossy_eval(m, in, out, n) * (T(1) - m.weight0);
    return d + g;
  }
  default: // Diffuse, DiffuseMirror (weight of the only continuous lobe is 1)
    return diffuse_eval(m);
  }
}

// reference: include/bsdf/union_bsdf.h:52-64, combined.h:82-96
template <class T> PT_INLINE T prob_continuous(const Material<T> &m) {
  switch (m.type) {
  case kMirror:
  case kDielectricRefractive:
    return T(0);
  case kDiffuseMirror:
    return m.weight0;
  default:
    return T(1);
  }
}

template <class T> struct DirSample {
  V3<T> dir;
  V3<T> multiplier;
  bool is_discrete;
};

// Cosine-weighted hemisphere about the normal facing the incoming ray.
// reference: include/bsdf/diffuse.h:24-46
template <class T, class R>
PT_INLINE DirSample<T> diffuse_sample(const Material<T> &m, V3<T> in, V3<T> n,
                                      R &rng) {
  T v0 = rng.next();
  T v1 = rng.next();
  V3<T> aligned = dot(in, n) > T(0) ? -n : n;
  V3<T> dir = find_relative_vec(aligned, v0, sqrt_(T(1) - v1), sqrt_(v1));
  return {dir, m.diffuse, false};
}

// reference: include/bsdf/glossy.h:35-51
template <class T, class R>
PT_INLINE DirSample<T> glossy_sample(const Material<T> &m, V3<T> in, V3<T> n,
                                     R &rng) {
  T v0 = rng.next();
  T v1 = rng.next();
  V3<T> refl = reflect_over_normal(in, n);
  T ct = Math<T>::pow(v1, m.inv_shininess_p1);
  T st = sqrt_(max_(T(0), T(1) - ct * ct));
  return {find_relative_vec(refl, v0, ct, st), m.specular, false};
}

template <class T, class R>
PT_INLINE DirSample<T> continuous_sample_bsdf(const Material<T> &m, V3<T> in,
                                              V3<T> n, R &rng) {
  switch (m.type) {
  case kGlossy:
    return glossy_sample(m, in, n, rng);
  case kDiffuseGlossy: {
    // reference: include/bsdf/combined.h:98-112 (lobe pick by inclusive search)
    T loc = rng.next();
    if (m.weight0 >= loc)
      return diffuse_sample(m, in, n, rng);
    return glossy_sample(m, in, n, rng);
  }
  case kDiffuseMirror: {
    (void)rng.next(); // Combined draws its lobe index even with one lobe
    return diffuse_sample(m, in, n, rng);
  }
  default:
    return diffuse_sample(m, in, n, rng);
  }
}

// Uniform sphere/hemisphere about the normal.
// reference: include/integrate/dir_sampler/uniform/uniform.h:24-40 and
// uniform_direction_sample.h:10-27
template <class T, class R>
PT_INLINE DirSample<T> continuous_sample_uniform(const Material<T> &m,
                                                 V3<T> in, V3<T> n, R &rng) {
  bool whole = !is_brdf(m);
  T v0 = rng.next();
  T v1 = rng.next();
  T ct = whole ? T(2) * v1 - T(1) : v1;
  T st = sqrt_(max_(T(0), T(1) - ct * ct));
  V3<T> dir = find_relative_vec(n, v0, ct, st);
  T inv_prob = whole ? T(12.566370614359172954) : T(6.283185307179586477);
  return {dir, (continuous_eval(m, in, dir, n) * dot(dir, n)) * inv_prob,
          false};
}

// reference: include/bsdf/mirror.h:18-21, dielectric_refractive.h:26-36,
// combined.h:114-128
template <class T, class R>
PT_INLINE DirSample<T> discrete_sample(const Material<T> &m, V3<T> in, V3<T> n,
                                       R &rng) {
  if (m.type == kDielectricRefractive) {
    T c = abs_(dot(in, n));
    T one_minus = T(1) - c;
    T prop = m.r0 + (T(1) - m.r0) * Math<T>::pow(one_minus, T(5));
    if (rng.next() < prop)
      return {reflect_over_normal(in, n), m.specular, true};
    return {refract_by_normal(m.ior, in, n), m.specular, true};
  }
  if (m.type == kDiffuseMirror)
    (void)rng.next(); // Combined lobe index draw
  return {reflect_over_normal(in, n), m.specular, true};
}

// Continuous-vs-discrete coin, then the chosen sampler.
// reference: include/integrate/dir_sampler/ref_from_continuous_sampler.h:13-43
template <class T, class R>
PT_INLINE DirSample<T> sample_direction(const IntegratorSettings &s,
                                        const Material<T> &m, V3<T> in,
                                        V3<T> n, R &rng) {
  T p = prob_continuous(m);
  bool discrete;
  if (p == T(0))
    discrete = true;
  else if (p == T(1))
    discrete = false;
  else
    discrete = p < rng.next();
  if (discrete)
    return discrete_sample(m, in, n, rng);
  if (s.dir_sampler == kDirUniform)
    return continuous_sample_uniform(m, in, n, rng);
  return continuous_sample_bsdf(m, in, n, rng);
}

// reference: include/integrate/term_prob/*/*.h
template <class T>
PT_INLINE T term_prob(const IntegratorSettings &s, uint32_t iters,
                      V3<T> mult) {
  switch (s.term_prob) {
  case kTermConstant:
    return T(s.constant_prob);
  case kTermNIters:
    return iters >= s.n_iters ? T(1) : T(0);
  case kTermNormalize: {
    T tp = T(1) - T(1) / (sum(mult) / T(3));
    return clamp_(tp, T(s.normalize_min_prob), T(1));
  }
  default: { // MultiplierFunc
    V3<T> one = {T(1), T(1), T(1)};
    V3<T> q = (mult / (mult + one)) * T(0.57);
    T sq = clamp_(squared_norm(q), T(0), T(1));
    T tp = abs_(Math<T>::pow(T(1) - sq, T(s.mf_exp)));
    return max_(tp, T(s.mf_min_prob));
  }
  }
}

// ---------------------------------------------------------------------------
// Scene view used during shading. Arrays are indexed by the flattened
// (world-space) triangle index carried in Hit::idx.
// ---------------------------------------------------------------------------

// Everything shading needs about one triangle in one record (six 16-byte
// vectors for float), so a hit costs a single dependent fetch.
template <class T> struct ShadeTri {
  V3<T> v0;
  uint32_t material;
  V3<T> e1;
  uint32_t pad0;
  V3<T> e2;
  uint32_t pad1;
  V3<T> n0;
  T pad2;
  V3<T> n1;
  T pad3;
  V3<T> n2;
  T pad4;
};

template <class T>
PT_HD inline ShadeTri<T> make_shade_tri(const TriGeom<T> &g, const V3<T> *n,
                                        uint32_t material) {
  ShadeTri<T> r{};
  r.v0 = g.v0;
  r.e1 = g.e1;
  r.e2 = g.e2;
  r.material = material;
  r.n0 = n[0];
  r.n1 = n[1];
  r.n2 = n[2];
  return r;
}

template <class T> struct SceneView {
  const ShadeTri<T> *tris; // [n_tris]
  const Material<T> *materials;
  const TriGeom<T> *light_tris; // emissive triangles, world space
  const T *light_cdf;           // inclusive cumulative weights, last = 1
  uint32_t n_lights;
};

template <class T> struct LightSample {
  V3<T> dir;
  T multiplier; // 1 / pdf of this direction (reference FSample::multiplier)
  T target_distance;
};

// reference: include/integrate/light_sampler/random_triangle/random_triangle.h:35-90
// and include/integrate/sample_triangle.h:10-28
template <class T, class R>
PT_INLINE bool sample_light(const SceneView<T> &sc, V3<T> position,
                            const Material<T> &mat, V3<T> normal, R &rng,
                            LightSample<T> *out) {
  if (sc.n_lights == 0)
    return false;
  T search = rng.next();
  uint32_t idx = 0;
  while (idx + 1 < sc.n_lights && !(sc.light_cdf[idx] >= search))
    ++idx;
  const TriGeom<T> tri = sc.light_tris[idx];
  T s = rng.next();
  T t = rng.next();
  if (s + t > T(1)) {
    s = T(1) - s;
    t = T(1) - t;
  }
  V3<T> point = baryo_to_point(tri, s, t);
  V3<T> dir_un = point - position;
  if (is_brdf(mat) && dot(dir_un, normal) <= T(0))
    return false;
  V3<T> dir = normalize(dir_un);
  V3<T> tri_n = normal_scaled_by_area(tri);
  T prob = idx == 0 ? sc.light_cdf[0] : sc.light_cdf[idx] - sc.light_cdf[idx - 1];
  T nw = abs_(dot(normal, dir) * dot(tri_n, dir));
  if (nw <= T(1e-8f)) // float nw < 1e-8 (double) <=> nw <= float(1e-8)
    return false;
  T sq = squared_norm(dir_un);
  T weight = nw / sq;
  out->dir = dir;
  out->multiplier = prob / weight;
  out->target_distance = sqrt_(sq);
  return true;
}

// ---------------------------------------------------------------------------
// Path state and one iteration.
// reference: include/integrate/rendering_equation_state.h:14-72 and
// include/integrate/rendering_equation_iteration.h:9-153
// ---------------------------------------------------------------------------
template <class T> struct PathState {
  uint32_t iters;
  uint32_t count_emission;
  uint32_t has_next;
  uint32_t n_light;       // 0 or 1 pending light samples
  uint32_t ray_has_target; // ray_info.target_distance has a value
  T ray_target;
  V3<T> ray_mult;   // ray_info.multiplier
  V3<T> light_mult; // light_samples[0].multiplier
  T light_target;   // light_samples[0].target_distance
  V3<T> total;      // float_rgb_total
};

template <class T> PT_INLINE PathState<T> initial_path_state() {
  PathState<T> s{};
  s.iters = 0;
  s.count_emission = 1;
  s.has_next = 1;
  s.n_light = 0;
  s.ray_has_target = 0;
  s.ray_target = T(0);
  s.ray_mult = {T(1), T(1), T(1)};
  s.total = {T(0), T(0), T(0)};
  return s;
}

// |target - t| < 1e-6 with the float difference compared against the double
// literal of the reference, i.e. <= float(1e-6) for float operands.
template <class T> PT_INLINE bool within_target(T target, T t) {
  return abs_(target - t) <= T(1e-6f);
}

// Surface data of the vertex shaded by one iteration (G-buffer output).
// What one iteration added, for derivative replays: the emission terms
// (material whose emission was scaled, and the multiplier it was scaled by)
// and the material of the vertex shaded (kNoMaterial if none).
constexpr uint32_t kNoMaterial = 0xffffffffu;
template <class T> struct IterTape {
  uint32_t n_emit;
  uint32_t emit_mat[2];
  V3<T> emit_mult[2];
  uint32_t vertex_mat;
};

template <class T> struct VertexInfo {
  V3<T> p;
  V3<T> normal;
  uint32_t material;
  uint32_t valid;
};

// Reflectance summary of a material for G-buffers: the diffuse albedo, the
// specular coefficient, or their lobe-weighted sum for Combined BSDFs.
template <class T> PT_INLINE V3<T> material_albedo(const Material<T> &m) {
  switch (m.type) {
  case kDiffuse:
    return m.diffuse;
  case kDiffuseGlossy:
  case kDiffuseMirror:
    return m.diffuse * m.weight0 + m.specular * (T(1) - m.weight0);
  default:
    return m.specular;
  }
}

// One vertex of the rendering equation. `light_hit` / `next_hit` are the
// closest hits of the light ray (when st.n_light) and of the continuation ray
// (when st.has_next) emitted by the previous step, with Hit::idx == kNoHit for
// a miss. Returns true when the path is finished (radiance in out->total).
// Otherwise writes the new state; out->n_light / out->has_next say which of
// *light_ray / *next_ray were emitted (both start at the new vertex).
template <class T, class R>
PT_INLINE bool rendering_equation_iteration(
    const PathState<T> &st, R &rng, const Ray<T> &last_ray,
    const Hit<T> &light_hit, const Hit<T> &next_hit,
    const IntegratorSettings &s, const SceneView<T> &sc, PathState<T> *out,
    Ray<T> *light_ray, Ray<T> *next_ray, VertexInfo<T> *vertex = nullptr,
    IterTape<T> *tape = nullptr) {
  PathState<T> ns;
  ns.iters = st.iters + 1;
  ns.total = st.total;
  ns.n_light = 0;
  ns.ray_has_target = 0;
  ns.ray_target = T(0);
  ns.ray_mult = st.ray_mult;
  ns.light_mult = {T(0), T(0), T(0)};
  ns.light_target = T(0);
  ns.count_emission = 0;
  ns.has_next = 0;

  if (tape) {
    tape->n_emit = 0;
    tape->vertex_mat = kNoMaterial;
  }
  if (st.n_light && light_hit.idx != kNoHit &&
      within_target(st.light_target, light_hit.t) &&
      !(s.back_cull_emission && light_hit.back)) {
    const uint32_t lm = sc.tris[light_hit.idx].material;
    ns.total += sc.materials[lm].emission * st.light_mult;
    if (tape) {
      tape->emit_mat[tape->n_emit] = lm;
      tape->emit_mult[tape->n_emit] = st.light_mult;
      tape->n_emit++;
    }
  }

  if (!st.has_next) {
    *out = ns;
    return true;
  }

  const Hit<T> &h = next_hit;
  if (h.idx == kNoHit ||
      (st.ray_has_target && !within_target(st.ray_target, h.t))) {
    *out = ns;
    return true;
  }

  V3<T> p = hit_point(last_ray, h.t);
  V3<T> mult = st.ray_mult;
  const ShadeTri<T> tri = sc.tris[h.idx];
  const Material<T> mat = sc.materials[tri.material];
  bool performs_samples = s.light_sampler != kLightNone;
  if ((!performs_samples || st.count_emission) &&
      !(s.back_cull_emission && h.back)) {
    ns.total += mult * mat.emission;
    if (tape) {
      tape->emit_mat[tape->n_emit] = tri.material;
      tape->emit_mult[tape->n_emit] = mult;
      tape->n_emit++;
    }
  }
  if (tape)
    tape->vertex_mat = tri.material;

  const TriGeom<T> g = {tri.v0, tri.e1, tri.e2};
  const V3<T> vn[3] = {tri.n0, tri.n1, tri.n2};
  V3<T> normal = shading_normal(g, vn, p);
  V3<T> in = last_ray.d;
  if (vertex)
    *vertex = {p, normal, tri.material, 1u};

  if (performs_samples) {
    T pc = prob_continuous(mat);
    if (pc > T(0)) {
      LightSample<T> ls;
      if (sample_light(sc, p, mat, normal, rng, &ls)) {
        ns.light_mult =
            ((mult * continuous_eval(mat, in, ls.dir, normal)) * pc) /
            ls.multiplier;
        ns.light_target = ls.target_distance;
        ns.n_light = 1;
        *light_ray = {p, ls.dir};
      }
    }
  }

  DirSample<T> ds = sample_direction(s, mat, in, normal, rng);
  ns.count_emission = ds.is_discrete ? 1u : 0u;
  V3<T> nm = mult * ds.multiplier;
  T tp = term_prob(s, st.iters, nm);
  ns.has_next = rng.next() > tp ? 1u : 0u;
  if (ns.has_next) {
    T denom = T(1) - tp;
    nm = {nm.x / denom, nm.y / denom, nm.z / denom};
    ns.ray_mult = nm;
    *next_ray = {p, ds.dir};
  } else if (!ns.n_light) {
    *out = ns;
    return true;
  }
  *out = ns;
  return false;
}

} // namespace pt
