// Property check of triangle_screen_rejects (core/geometry.h) on the host:
// near-boundary rays (u ~ 0, u ~ 1, v ~ 0, u + v ~ 1, t ~ 1e-5, t ~ t_best)
// against random triangles, with the reciprocal of a perturbed by up to
// `max_ulps` ulps around the correctly rounded 1/a. A violation is a triangle
// that the exact test accepts with t <= t_best but the screen rejects.
#include <cmath>
#include <cstdint>
#include <random>

#include "../../path_amd/csrc/core/geometry.h"

using namespace pt;

static float step_ulps(float x, int k) {
  for (; k > 0; --k)
    x = std::nextafter(x, INFINITY);
  for (; k < 0; ++k)
    x = std::nextafter(x, -INFINITY);
  return x;
}

extern "C" int64_t screen_violations(uint64_t seed, int64_t n, float margin, int max_ulps,
                                     int64_t *n_rejected, int64_t *n_accepted) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<float> U(-1.f, 1.f), P(0.f, 1.f);
  int64_t bad = 0, rej = 0, acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    const V3<float> v0 = {U(g), U(g), U(g)}, e1 = {U(g), U(g), U(g)}, e2 = {U(g), U(g), U(g)};
    // barycentric target near one of the edges or inside
    const float eps = std::ldexp(U(g), -int(g() % 30));
    float u = P(g), v = P(g) * (1.f - u);
    switch (g() % 5) {
    case 0: u = eps; break;
    case 1: u = 1.f + eps; v = eps * 0.5f; break;
    case 2: v = eps; break;
    case 3: v = 1.f - u + eps; break;
    default: break;
    }
    const V3<float> p = v0 + e1 * u + e2 * v;
    V3<float> d = {U(g), U(g), U(g)};
    float dist = (g() % 3 == 0) ? 1e-5f * (1.f + std::ldexp(U(g), -int(g() % 24)))
                                : std::ldexp(P(g) + 0.01f, int(g() % 8) - 4);
    const Ray<float> ray = {p - d * dist, d};
    const TriGeom<float> tri = {v0, e1, e2};
    const V3<float> h = cross(ray.d, tri.e2);
    const float a = dot(tri.e1, h);
    const V3<float> s = ray.o - tri.v0;
    const float x = dot(s, h);
    const V3<float> q = cross(s, tri.e1);
    const float y = dot(ray.d, q);
    const float z = dot(tri.e2, q);
    float t = 0.f;
    uint32_t back = 0;
    const bool hit = intersect_triangle_exact(a, x, y, z, &t, &back);
    // the exact test above must be intersect_triangle itself
    float t2 = 0.f;
    uint32_t back2 = 0;
    const bool hit2 = intersect_triangle(ray, tri, &t2, &back2);
    if (hit != hit2 || (hit && (t != t2 || back != back2)))
      return -1;
    // t_best at, just above or just below the hit distance, or unbounded
    float t_best = 3.4028235e38f;
    if (hit) {
      switch (g() % 4) {
      case 0: t_best = t; break;
      case 1: t_best = step_ulps(t, 1 + int(g() % 4)); break;
      case 2: t_best = step_ulps(t, -1 - int(g() % 4)); break;
      default: break;
      }
    }
    const float f = 1.0f / a;
    for (int k = -max_ulps; k <= max_ulps; ++k) {
      const float r = step_ulps(f, k);
      const bool rejects = triangle_screen_rejects(a, x, y, z, r, t_best, margin);
      rej += rejects;
      if (hit && t <= t_best) {
        ++acc;
        bad += rejects;
      }
    }
  }
  *n_rejected = rej;
  *n_accepted = acc;
  return bad;
}
