// BVH traversal on gfx950: one ray per lane, near-first order, per-lane stack.
//
// Node layout (64 B, both child boxes in the parent) and leaf-ordered triangles
// are produced by host/bvh.cpp. Child refs >= 0 are inner nodes, refs < 0 are
// ~(first triangle of a leaf); the last triangle of a leaf has bit 31 of its
// index word set.
//
// The closest hit is the minimum of (t, flattened triangle index) over all
// triangles the float Moller-Trumbore test accepts. Boxes are padded on the
// host and a child is skipped only when its entry distance is strictly beyond
// the current best distance, so the result is independent of the visiting
// order (and equal to a brute-force loop over all triangles).
#pragma once

#include <hip/hip_runtime.h>

#include "../core/geometry.h"
#include "kernels.h"

namespace pt {
namespace gpu {

constexpr uint32_t kLeafEnd = 0x80000000u;

// Stack in private (scratch) memory: used where a kernel has no LDS budget.
struct PrivateStack {
  int v[kTraceStackSize];
  int sp = 0;
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ void push(int x) { v[sp++] = x; }
  __device__ __forceinline__ int pop() { return v[--sp]; }
};

// The top kLdsStack entries live in LDS (one column per lane of a 256-thread
// block); deeper entries rotate into a per-thread global overflow area, laid
// out entry-major so a wave's spills are contiguous. Logical entry k sits at
// LDS row k % kLdsStack while k >= sp - kLdsStack.
constexpr int kLdsStack = 16;
constexpr int kOverflowDepth = kTraceStackSize - kLdsStack;

struct LdsStack {
  int *lds;       // &shared[threadIdx.x]; row stride 256
  int *overflow;  // &global[thread id]; entry stride `ostride`
  uint32_t ostride;
  int sp = 0;
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ void push(int x) {
    int *slot = lds + 256 * (sp & (kLdsStack - 1));
    if (sp >= kLdsStack)
      overflow[size_t(sp - kLdsStack) * ostride] = *slot;
    *slot = x;
    ++sp;
  }
  __device__ __forceinline__ int pop() {
    --sp;
    int *slot = lds + 256 * (sp & (kLdsStack - 1));
    int x = *slot;
    if (sp >= kLdsStack)
      *slot = overflow[size_t(sp - kLdsStack) * ostride];
    return x;
  }
};

// Returns the closest hit with t <= t_limit. If occlude_below > 0, stops at
// the first hit closer than occlude_below (any such hit makes a shadow ray's
// light sample fail its distance test, so the caller cannot tell the
// difference).
template <class Stack>
__device__ __forceinline__ Hit<float>
trace_closest(const float4 *__restrict__ nodes, const float4 *__restrict__ tris,
              const Ray<float> &ray, float t_limit, float occlude_below,
              Stack &stack) {
  V3<float> inv = inv_direction(ray.d);
  V3<float> oi = {ray.o.x * inv.x, ray.o.y * inv.y, ray.o.z * inv.z};
  float best_t = t_limit;
  uint32_t best_idx = kNoHit;
  uint32_t best_back = 0;
  int node = 0;
  for (;;) {
    if (node >= 0) {
      const float4 *np = nodes + 4 * node;
      float4 a = np[0], b = np[1], c = np[2];
      int4 d = __builtin_bit_cast(int4, np[3]);
      float lx0 = fmaf(a.x, inv.x, -oi.x), lx1 = fmaf(a.y, inv.x, -oi.x);
      float ly0 = fmaf(a.z, inv.y, -oi.y), ly1 = fmaf(a.w, inv.y, -oi.y);
      float lz0 = fmaf(c.x, inv.z, -oi.z), lz1 = fmaf(c.y, inv.z, -oi.z);
      float rx0 = fmaf(b.x, inv.x, -oi.x), rx1 = fmaf(b.y, inv.x, -oi.x);
      float ry0 = fmaf(b.z, inv.y, -oi.y), ry1 = fmaf(b.w, inv.y, -oi.y);
      float rz0 = fmaf(c.z, inv.z, -oi.z), rz1 = fmaf(c.w, inv.z, -oi.z);
      float lmin = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)),
                         fmaxf(fminf(lz0, lz1), 0.0f));
      float lmax = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)),
                         fminf(fmaxf(lz0, lz1), best_t));
      float rmin = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)),
                         fmaxf(fminf(rz0, rz1), 0.0f));
      float rmax = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)),
                         fminf(fmaxf(rz0, rz1), best_t));
      bool hl = lmin <= lmax;
      bool hr = rmin <= rmax;
      if (hl && hr) {
        int near = d.x, far = d.y;
        if (rmin < lmin) {
          near = d.y;
          far = d.x;
        }
        stack.push(far);
        node = near;
      } else if (hl) {
        node = d.x;
      } else if (hr) {
        node = d.y;
      } else {
        if (stack.empty())
          break;
        node = stack.pop();
      }
      continue;
    }
    // leaf
    uint32_t ti = uint32_t(~node);
    for (;;) {
      const float4 *tp = tris + 3 * ti;
      float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
      uint32_t word = __float_as_uint(p0.w);
      uint32_t idx = word & ~kLeafEnd;
      TriGeom<float> tri = {{p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z},
                            {p2.x, p2.y, p2.z}};
      float t;
      uint32_t back;
      if (intersect_triangle(ray, tri, &t, &back) &&
          hit_before(t, idx, best_t, best_idx)) {
        best_t = t;
        best_idx = idx;
        best_back = back;
      }
      if (word & kLeafEnd)
        break;
      ++ti;
    }
    if (best_idx != kNoHit && best_t < occlude_below)
      break;
    if (stack.empty())
      break;
    node = stack.pop();
  }
  return {best_t, best_idx, best_back};
}

// Hit record packed in 8 bytes: t bits, index with the back-face flag in
// bit 31 (triangle indices stay below 2^31 - 1), 0xffffffff for a miss.
__device__ __forceinline__ uint2 pack_hit(const Hit<float> &h) {
  return make_uint2(__float_as_uint(h.t),
                    h.idx == kNoHit ? kNoHit : (h.idx | (h.back << 31)));
}
__device__ __forceinline__ Hit<float> unpack_hit(uint2 v) {
  if (v.y == kNoHit)
    return {__uint_as_float(v.x), kNoHit, 0};
  return {__uint_as_float(v.x), v.y & 0x7fffffffu, v.y >> 31};
}

} // namespace gpu
} // namespace pt
