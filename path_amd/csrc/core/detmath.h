// Deterministic single-precision transcendental functions.
//
// The GPU library's sinf/powf and the host libm's differ in their last bits,
// which would make a float render on the GPU and the same float render on the
// host drift apart path by path. These functions are built only from IEEE
// operations that are correctly rounded on both sides (+, -, *, /, fma, floor
// and bit manipulation), so Math<float> returns the same bits on gfx950 and on
// x86-64. Method: exact range reduction followed by truncated Taylor series
// (coefficients are (2*pi)^k/k!, ln2^k/k! and 2/((2k+1) ln 2), rounded to
// float; exact values in the comments). Accuracy over the ranges used by the
// renderer is a few ulp, checked against libm in double by the test suite.
//
// Math<double> forwards to libm and is used only by the FP64 CPU checker.
#pragma once

#include "pt_defs.h"

#include <cmath>

namespace pt {

PT_INLINE uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
PT_INLINE float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

namespace detmath {

// sin(2*pi*v), cos(2*pi*v) for |v| < 2^21.
PT_INLINE void sincos_2pi(float v, float *s_out, float *c_out) {
  float q = __builtin_floorf(v * 4.0f + 0.5f); // nearest quarter turn
  float r = v - q * 0.25f;                     // |r| <= 1/8, exact
  int quadrant = int(q) & 3;
  float r2 = r * r;
  // sin(2 pi r) = r * sum_k (-1)^k (2pi)^(2k+1) / (2k+1)! r^(2k)
  float s = 0x1.e8f434p+1f;
  s = fma_(s, r2, -0x1.e3075p+3f);
  s = fma_(s, r2, 0x1.507834p+5f);
  s = fma_(s, r2, -0x1.32d2ccp+6f);
  s = fma_(s, r2, 0x1.466bc6p+6f);
  s = fma_(s, r2, -0x1.4abbcep+5f);
  s = fma_(s, r2, 0x1.921fb6p+2f);
  s = s * r;
  // cos(2 pi r) = sum_k (-1)^k (2pi)^(2k) / (2k)! r^(2k)
  float c = -0x1.b6e25p+0f;
  c = fma_(c, r2, 0x1.f9d38ap+2f);
  c = fma_(c, r2, -0x1.a6d1f2p+4f);
  c = fma_(c, r2, 0x1.e1f506p+5f);
  c = fma_(c, r2, -0x1.55d3c8p+6f);
  c = fma_(c, r2, 0x1.03c1f0p+6f);
  c = fma_(c, r2, -0x1.3bd3ccp+4f);
  c = fma_(c, r2, 1.0f);
  float so, co;
  switch (quadrant) {
  case 0:
    so = s;
    co = c;
    break;
  case 1:
    so = c;
    co = -s;
    break;
  case 2:
    so = -s;
    co = -c;
    break;
  default:
    so = -c;
    co = s;
    break;
  }
  *s_out = so;
  *c_out = co;
}

// log2(x) for finite x > 0.
PT_INLINE float log2(float x) {
  int e_adj = 0;
  if (x < 0x1p-126f) { // subnormal input: rescale
    x = x * 0x1p+24f;
    e_adj = -24;
  }
  uint32_t bits = f2u(x);
  int e = int((bits >> 23) & 0xffu) - 127 + e_adj;
  float m = u2f((bits & 0x7fffffu) | 0x3f800000u); // [1, 2)
  if (m > 0x1.6a09e6p+0f) {                         // sqrt(2)
    m = m * 0.5f;
    e += 1;
  }
  // log2(m) = (2 / ln 2) * atanh(z), z = (m - 1) / (m + 1), |z| < 0.1716
  float z = (m - 1.0f) / (m + 1.0f);
  float z2 = z * z;
  float p = 0x1.c68f56p-3f;
  p = fma_(p, z2, 0x1.0c9a84p-2f);
  p = fma_(p, z2, 0x1.484b14p-2f);
  p = fma_(p, z2, 0x1.a61762p-2f);
  p = fma_(p, z2, 0x1.2776c6p-1f);
  p = fma_(p, z2, 0x1.ec709ep-1f);
  p = fma_(p, z2, 0x1.715476p+1f);
  return fma_(p, z, float(e));
}

// 2^x.
PT_INLINE float exp2(float x) {
  if (x < -152.0f)
    return 0.0f;
  if (x > 128.0f)
    return u2f(0x7f800000u);
  float n = __builtin_floorf(x + 0.5f);
  float f = x - n; // [-0.5, 0.5]
  // 2^f = sum_k (ln 2)^k / k! f^k
  float p = 0x1.e4cf52p-28f;
  p = fma_(p, f, 0x1.b5253ep-24f);
  p = fma_(p, f, 0x1.62c022p-20f);
  p = fma_(p, f, 0x1.ffcbfcp-17f);
  p = fma_(p, f, 0x1.430912p-13f);
  p = fma_(p, f, 0x1.5d87fep-10f);
  p = fma_(p, f, 0x1.3b2ab6p-7f);
  p = fma_(p, f, 0x1.c6b08ep-5f);
  p = fma_(p, f, 0x1.ebfbe0p-3f);
  p = fma_(p, f, 0x1.62e430p-1f);
  p = fma_(p, f, 1.0f);
  int ni = int(n);
  if (ni >= -126) {
    return p * u2f(uint32_t(ni + 127) << 23);
  }
  // two-step scaling into the subnormal range
  return (p * u2f(uint32_t(ni + 127 + 64) << 23)) * 0x1p-64f;
}

// x^y for x >= 0 (the renderer never raises a negative base).
PT_INLINE float pow(float x, float y) {
  if (x == 0.0f)
    return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : u2f(0x7f800000u));
  if (x == 1.0f || y == 0.0f)
    return 1.0f;
  return exp2(y * log2(x));
}

} // namespace detmath

template <class T> struct Math;

template <> struct Math<float> {
  static PT_INLINE void sincos_2pi(float v, float *s, float *c) {
    detmath::sincos_2pi(v, s, c);
  }
  static PT_INLINE float pow(float x, float y) { return detmath::pow(x, y); }
};

template <> struct Math<double> {
  static inline void sincos_2pi(double v, double *s, double *c) {
    const double two_pi = 6.283185307179586476925286766559;
    *s = std::sin(two_pi * v);
    *c = std::cos(two_pi * v);
  }
  static inline double pow(double x, double y) { return std::pow(x, y); }
};

} // namespace pt
