// Pinhole camera ray for a film position.
// reference: include/render/detail/integrate_image/initial_ray_sample.h:16-31
#pragma once

#include "geometry.h"

namespace pt {

// film_to_world is row-major 3x4 (p' = M p + t); the ray starts at t.
template <class T>
PT_INLINE Ray<T> camera_ray(const T *m, T x, T y, uint32_t x_dim,
                            uint32_t y_dim) {
  T cx = (T(2) * x) / T(x_dim) - T(1);
  T cy = (T(-2) * y) / T(y_dim) + T(1);
  T cz = T(-1);
  V3<T> o = {m[3], m[7], m[11]};
  V3<T> w = {fma_(m[2], cz, fma_(m[1], cy, m[0] * cx)) + m[3],
             fma_(m[6], cz, fma_(m[5], cy, m[4] * cx)) + m[7],
             fma_(m[10], cz, fma_(m[9], cy, m[8] * cx)) + m[11]};
  return {o, normalize(w - o)};
}

// FloatRGB -> BGRA32 with the x / (x + 1) tone map, truncating conversion.
// reference: include/lib/bgra_32.h:12-31, include/lib/tone_map.h:7-10
PT_INLINE uint32_t float_rgb_to_bgra32(float r, float g, float b) {
  float tr = r / (r + 1.0f), tg = g / (g + 1.0f), tb = b / (b + 1.0f);
  uint32_t ur = uint32_t(tr * 255.0f), ug = uint32_t(tg * 255.0f),
           ub = uint32_t(tb * 255.0f);
  // memory order B, G, R, A(=0)
  return (ub & 0xffu) | ((ug & 0xffu) << 8) | ((ur & 0xffu) << 16);
}

} // namespace pt
