// Ray / triangle / box primitives shared by the GPU kernels and the CPU checker.
#pragma once

#include "vec3.h"

namespace pt {

template <class T> struct Ray {
  V3<T> o;
  V3<T> d; // unit length
};

// Closest-hit record. `idx` is the triangle's index in the flattened scene
// (world-space triangle order), independent of any acceleration layout.
template <class T> struct Hit {
  T t;
  uint32_t idx;
  uint32_t back; // 1 if the ray hit the triangle's back side
};

constexpr uint32_t kNoHit = 0xffffffffu;

// Triangle stored as v0 and the two edges e1 = v1 - v0, e2 = v2 - v0 (the same
// float subtractions the reference performs in every intersection test).
template <class T> struct TriGeom {
  V3<T> v0, e1, e2;
};

// Moller-Trumbore. reference: include/intersect/triangle_impl.h:8-39
// (|a| < 1e-7 rejects, u in [0,1], v >= 0, u + v <= 1, hit iff t > 1e-5,
// back face iff a < 0). The double-precision literals of the reference compare
// against float operands; for floats `x < 1e-7` is the same predicate as
// `x < 1e-7f` and `x > 1e-5` the same as `x > 1e-5f`.
template <class T>
PT_INLINE bool intersect_triangle(const Ray<T> &ray, const TriGeom<T> &tri,
                                  T *t_out, uint32_t *back_out) {
  V3<T> h = cross(ray.d, tri.e2);
  T a = dot(tri.e1, h);
  if (abs_(a) < T(1e-7f))
    return false;
  T f = T(1) / a;
  V3<T> s = ray.o - tri.v0;
  T u = f * dot(s, h);
  if (u < T(0) || u > T(1))
    return false;
  V3<T> q = cross(s, tri.e1);
  T v = f * dot(ray.d, q);
  if (v < T(0) || u + v > T(1))
    return false;
  T t = f * dot(tri.e2, q);
  if (t > T(1e-5f)) {
    *t_out = t;
    *back_out = a < T(0) ? 1u : 0u;
    return true;
  }
  return false;
}

// Screen in front of intersect_triangle for traversal loops. With
//   a = dot(e1, h), x = dot(s, h), y = dot(d, q), z = dot(e2, q)
// (the quantities of intersect_triangle) and r any approximation of 1/a with
// relative error below 2^-21 (a hardware reciprocal), the screen evaluates
// u' = r x, v' = r y, t' = r z and returns true only when its margin makes
// the exact test reject the triangle or find t > t_best:
//   |a| < 1e-7                 (the exact test's own first check)
//   u' < -2^-100            => u < 0 (same sign, no underflow to zero)
//   u' > 1 + m              => u > 1
//   v' < -2^-100            => v < 0
//   u' + v' > 1 + m         => fl(u + v) > 1 (u', v' >= -2^-100 here)
//   t' < 1e-5 (1 - m)       => t < 1e-5
//   t' > t_best (1 + m)     => t > t_best
// for m = 2^-16 (the product errors stay below 2^-20 relative). NaN fails
// every comparison, so such a triangle reaches the exact test.
template <class T>
PT_INLINE bool triangle_screen_rejects(T a, T x, T y, T z, T r, T t_best,
                                       T m = T(0x1p-16f)) {
  const T u = r * x, v = r * y, t = r * z;
  const T tiny = T(0x1p-100f), one = T(1) + m;
  return abs_(a) < T(1e-7f) || u < -tiny || u > one || v < -tiny || u + v > one ||
         t < T(1e-5f) * (T(1) - m) || t > t_best * one;
}

// intersect_triangle on the screen's quantities: the same operations
// (f = 1 / a correctly rounded, then f * x, f * y, f * z) and predicates.
template <class T>
PT_INLINE bool intersect_triangle_exact(T a, T x, T y, T z, T *t_out,
                                        uint32_t *back_out) {
  if (abs_(a) < T(1e-7f))
    return false;
  const T f = T(1) / a;
  const T u = f * x;
  if (u < T(0) || u > T(1))
    return false;
  const T v = f * y;
  if (v < T(0) || u + v > T(1))
    return false;
  const T t = f * z;
  if (t > T(1e-5f)) {
    *t_out = t;
    *back_out = a < T(0) ? 1u : 0u;
    return true;
  }
  return false;
}

// Total order used to pick the closest hit: smaller distance first, then the
// smaller flattened triangle index. The reference keeps whichever equal-distance
// candidate it tested last (include/intersect/optional_min.h:16-25), which
// depends on traversal order; this rule makes the closest hit a function of the
// ray and the triangle set alone.
template <class T>
PT_INLINE bool hit_before(T t, uint32_t idx, T best_t, uint32_t best_idx) {
  return t < best_t || (t == best_t && idx < best_idx);
}

// reference: include/intersect/intersection.h:17-20
template <class T> PT_INLINE V3<T> hit_point(const Ray<T> &ray, T t) {
  return ray.o + ray.d * t;
}

// Barycentric weights of a point. reference: include/intersect/triangle_impl.h:55-72
template <class T>
PT_INLINE void interpolation_values(const TriGeom<T> &tri, V3<T> p, T *w0,
                                    T *w1, T *w2) {
  V3<T> p2 = p - tri.v0;
  T d00 = dot(tri.e1, tri.e1);
  T d01 = dot(tri.e1, tri.e2);
  T d11 = dot(tri.e2, tri.e2);
  T d20 = dot(p2, tri.e1);
  T d21 = dot(p2, tri.e2);
  T denom = fma_(d00, d11, -(d01 * d01));
  T v = fma_(d11, d20, -(d01 * d21)) / denom;
  T w = fma_(d00, d21, -(d01 * d20)) / denom;
  *w0 = (T(1) - v) - w;
  *w1 = v;
  *w2 = w;
}

// Shading normal: normalized barycentric blend of the vertex normals.
// reference: include/scene/triangle_data.h:14-20
template <class T>
PT_INLINE V3<T> shading_normal(const TriGeom<T> &tri, const V3<T> n[3],
                               V3<T> p) {
  T w0, w1, w2;
  interpolation_values(tri, p, &w0, &w1, &w2);
  return normalize((n[0] * w0 + n[1] * w1) + n[2] * w2);
}

// reference: include/intersect/triangle.h:49-51 (0.5 * cross(e1, e2))
template <class T> PT_INLINE V3<T> normal_scaled_by_area(const TriGeom<T> &tri) {
  return cross(tri.e1, tri.e2) * T(0.5);
}

// reference: include/intersect/triangle.h:73-79
template <class T>
PT_INLINE V3<T> baryo_to_point(const TriGeom<T> &tri, T s, T t) {
  return (tri.e1 * s + tri.e2 * t) + tri.v0;
}

// Zero direction components become 1e-20 before inversion.
// reference: include/intersect/accel/aabb.h:16-28
template <class T> PT_INLINE V3<T> inv_direction(V3<T> d) {
  if (d.x == T(0))
    d.x = T(1e-20f);
  if (d.y == T(0))
    d.y = T(1e-20f);
  if (d.z == T(0))
    d.z = T(1e-20f);
  return {T(1) / d.x, T(1) / d.y, T(1) / d.z};
}

// Slab test, no t >= 0 clamp. reference: include/intersect/accel/aabb.h:112-131
template <class T>
PT_INLINE bool slab_test(V3<T> lo, V3<T> hi, V3<T> o, V3<T> inv_d, T *tmin_out,
                         T *tmax_out) {
  T tx0 = (lo.x - o.x) * inv_d.x, tx1 = (hi.x - o.x) * inv_d.x;
  T ty0 = (lo.y - o.y) * inv_d.y, ty1 = (hi.y - o.y) * inv_d.y;
  T tz0 = (lo.z - o.z) * inv_d.z, tz1 = (hi.z - o.z) * inv_d.z;
  T tmin = max_(max_(min_(tx0, tx1), min_(ty0, ty1)), min_(tz0, tz1));
  T tmax = min_(min_(max_(tx0, tx1), max_(ty0, ty1)), max_(tz0, tz1));
  *tmin_out = tmin;
  *tmax_out = tmax;
  return tmin <= tmax;
}

} // namespace pt
