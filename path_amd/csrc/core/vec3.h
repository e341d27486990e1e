// Small fixed-size vector used by all shared geometry/shading code.
// Operation order is part of the contract (see pt_defs.h): dot products are
// evaluated as fma(z, z', fma(y, y', x * x')), cross products as
// fma(a, b, -(c * d)).
#pragma once

#include "pt_defs.h"

namespace pt {

template <class T> struct V3 {
  T x, y, z;
};

using float3_ = V3<float>;

template <class T> PT_INLINE V3<T> v3(T x, T y, T z) { return {x, y, z}; }
template <class T> PT_INLINE V3<T> operator+(V3<T> a, V3<T> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T> PT_INLINE V3<T> operator-(V3<T> a, V3<T> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T> PT_INLINE V3<T> operator-(V3<T> a) {
  return {-a.x, -a.y, -a.z};
}
template <class T> PT_INLINE V3<T> operator*(V3<T> a, T s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class T> PT_INLINE V3<T> operator*(T s, V3<T> a) {
  return {a.x * s, a.y * s, a.z * s};
}
// Component-wise product (FloatRGB arithmetic).
template <class T> PT_INLINE V3<T> operator*(V3<T> a, V3<T> b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class T> PT_INLINE V3<T> operator/(V3<T> a, V3<T> b) {
  return {a.x / b.x, a.y / b.y, a.z / b.z};
}
template <class T> PT_INLINE V3<T> operator/(V3<T> a, T s) {
  return {a.x / s, a.y / s, a.z / s};
}
template <class T> PT_INLINE V3<T> &operator+=(V3<T> &a, V3<T> b) {
  a = a + b;
  return a;
}

template <class T> PT_INLINE T dot(V3<T> a, V3<T> b) {
  return fma_(a.z, b.z, fma_(a.y, b.y, a.x * b.x));
}
template <class T> PT_INLINE V3<T> cross(V3<T> a, V3<T> b) {
  return {fma_(a.y, b.z, -(a.z * b.y)), fma_(a.z, b.x, -(a.x * b.z)),
          fma_(a.x, b.y, -(a.y * b.x))};
}
template <class T> PT_INLINE T squared_norm(V3<T> a) { return dot(a, a); }
template <class T> PT_INLINE T sum(V3<T> a) { return (a.x + a.y) + a.z; }

// Eigen-compatible normalization: a zero vector is returned unchanged.
// reference: Eigen MatrixBase::normalized (used through UnitVector::new_normalize,
// include/lib/unit_vector.h).
template <class T> PT_INLINE V3<T> normalize(V3<T> a) {
  T sq = squared_norm(a);
  if (sq > T(0)) {
    T inv = T(1) / sqrt_(sq);
    return a * inv;
  }
  return a;
}

template <class T, class U> PT_INLINE V3<T> cast3(V3<U> a) {
  return {T(a.x), T(a.y), T(a.z)};
}

} // namespace pt
