// Common definitions for code compiled both by hipcc (device + host) and by a
// plain host C++ compiler (the CPU checker and host tools).
//
// Everything under core/ is header-only, templated on the scalar type T
// (float on the GPU, float or double on the host) and written so that the
// float instantiation performs the same IEEE-754 operations in the same order
// on the host and on gfx950: only +, -, *, /, sqrt and explicit fused
// multiply-add are used, and every translation unit is compiled with
// -ffp-contract=off so the compiler never fuses or reorders on its own.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define PT_HD __host__ __device__
#define PT_INLINE __host__ __device__ inline __attribute__((always_inline))
#else
#define PT_HD
#define PT_INLINE inline __attribute__((always_inline))
#endif

namespace pt {

template <class T> PT_INLINE T fma_(T a, T b, T c);
template <> PT_INLINE float fma_<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <> PT_INLINE double fma_<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

template <class T> PT_INLINE T sqrt_(T a);
template <> PT_INLINE float sqrt_<float>(float a) { return __builtin_sqrtf(a); }
template <> PT_INLINE double sqrt_<double>(double a) {
  return __builtin_sqrt(a);
}

template <class T> PT_INLINE T abs_(T a) { return a < T(0) ? -a : a; }
template <class T> PT_INLINE T min_(T a, T b) { return b < a ? b : a; }
template <class T> PT_INLINE T max_(T a, T b) { return a < b ? b : a; }
template <class T> PT_INLINE T clamp_(T v, T lo, T hi) {
  return v < lo ? lo : (hi < v ? hi : v);
}

} // namespace pt
