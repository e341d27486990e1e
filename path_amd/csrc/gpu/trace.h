// BVH traversal on gfx950: one ray per lane, near-first order, per-lane stack.
//
// 4-wide nodes (64 B: grid origin and power-of-two steps, child refs, 8-bit quantised
// child boxes) and leaf-ordered triangles are produced by host/bvh.cpp. Child refs >= 0 are
// inner nodes, refs < 0 are ~(first triangle of a leaf), INT32_MIN is an empty
// slot; the last triangle of a leaf has bit 31 of its index word set.
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
constexpr int kEmptyChildRef = INT32_MIN; // host/bvh.h kEmptyChild
constexpr size_t kNodeFloat4 = 4;         // 64-byte nodes

// Stack in private (scratch) memory: used where a kernel has no LDS budget.
struct PrivateStack {
  int v[kTraceStackSize];
  int sp = 0;
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ void push(int x) { v[sp++] = x; }
  __device__ __forceinline__ int pop() { return v[--sp]; }
};

// The top kLdsStack entries live in LDS, deeper entries rotate into an
// overflow area (a per-thread column of a global buffer, laid out entry-major
// so a wave's spills are contiguous, or a private array). Each wave owns a
// kLdsStack x 64 tile of LDS, one column per lane (row stride 64 ints: the
// lanes of a row hit distinct banks). Logical entry k sits at LDS row
// k % kLdsStack while k >= sp - kLdsStack.
constexpr int kLdsStack = 16;
constexpr int kOverflowDepth = kTraceStackSize - kLdsStack;
constexpr uint32_t kLdsStackInts = kLdsStack * 64; // per wave

// This thread's LDS column inside `lds` (kLdsStackInts per wave of the block).
__device__ __forceinline__ int *lds_stack_column(int *lds) {
  return lds + (threadIdx.x >> 6) * kLdsStackInts + (threadIdx.x & 63u);
}

struct LdsStack {
  int *lds;       // lds_stack_column()
  int *overflow;  // entry k >= kLdsStack at overflow[(k - kLdsStack) * ostride]
  uint32_t ostride;
  int sp = 0;
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ void push(int x) {
    int *slot = lds + 64 * (sp & (kLdsStack - 1));
    if (sp >= kLdsStack)
      overflow[size_t(sp - kLdsStack) * ostride] = *slot;
    *slot = x;
    ++sp;
  }
  __device__ __forceinline__ int pop() {
    --sp;
    int *slot = lds + 64 * (sp & (kLdsStack - 1));
    int x = *slot;
    if (sp >= kLdsStack)
      *slot = overflow[size_t(sp - kLdsStack) * ostride];
    return x;
  }
};

// intersect_triangle with triangle_screen_rejects (core/geometry.h) in front,
// on the hardware reciprocal (v_rcp_f32, 1 ulp): most triangles are rejected
// without the correctly rounded division and the exact test's branches; the
// result equals intersect_triangle's whenever that one finds a hit with
// t <= t_best.
__device__ __forceinline__ bool intersect_triangle_screened(const Ray<float> &ray,
                                                            const TriGeom<float> &tri,
                                                            float t_best, float *t_out,
                                                            uint32_t *back_out) {
  const V3<float> h = cross(ray.d, tri.e2);
  const float a = dot(tri.e1, h);
  const V3<float> s = ray.o - tri.v0;
  const float x = dot(s, h);
  const V3<float> q = cross(s, tri.e1);
  const float y = dot(ray.d, q);
  const float z = dot(tri.e2, q);
  if (triangle_screen_rejects(a, x, y, z, __builtin_amdgcn_rcpf(a), t_best))
    return false;
  return intersect_triangle_exact(a, x, y, z, t_out, back_out);
}

// Resumable closest-hit traversal of one ray: init() then step() until it
// returns true. One step visits one inner node or one leaf, so a caller can
// interleave rays (lane refill). The closest hit is kept with t <= t_limit;
// if occlude_below > 0 the traversal stops at the first hit closer than
// occlude_below (any such hit makes a shadow ray's light sample fail its
// distance test, so the caller cannot tell the difference).
struct TraceState {
  Ray<float> ray;
  V3<float> inv, oi;
  float best_t, occlude_below;
  uint32_t best_idx, best_back;
  int node, pending;
  static constexpr int kNoRef = INT32_MIN; // no node / no held leaf

  __device__ __forceinline__ void init(const Ray<float> &r, float t_limit, float occl) {
    ray = r;
    inv = inv_direction(r.d);
    oi = {r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z};
    best_t = t_limit;
    occlude_below = occl;
    best_idx = kNoHit;
    best_back = 0;
    node = 0;
    pending = kNoRef;
  }

  __device__ __forceinline__ Hit<float> hit() const { return {best_t, best_idx, best_back}; }

  // Visit the inner node `node`: continue with its nearest hit child, push
  // the other hit children (far first), or pop when none is hit (kNoRef when
  // the stack is empty). Hit children are visited nearest first, also for
  // occlusion rays (visiting them in slot order measured 1% slower on the
  // 82k-triangle scene).
  template <class Stack>
  __device__ __forceinline__ void visit_inner(const float4 *__restrict__ nodes, Stack &stack) {
    // origin xyz, step x | step yz, child 0-1 | child 2-3, lo_x, hi_x |
    // lo_y, hi_y, lo_z, hi_z
    const float4 *np = nodes + kNodeFloat4 * size_t(node);
    const float4 f0 = np[0], f1 = np[1];
    const uint4 f2 = __builtin_bit_cast(uint4, np[2]), f3 = __builtin_bit_cast(uint4, np[3]);
    const uint4 qa = make_uint4(f2.z, f2.w, f3.x, f3.y); // lo_x, hi_x, lo_y, hi_y
    const uint2 qb = make_uint2(f3.z, f3.w);             // lo_z, hi_z
    // plane of code q on axis a: fma(q, step_a, origin_a); in ray distance
    // t = (plane - o) / d  ~  fma(q, step_a / d_a, (origin_a - o_a) / d_a)
    const float sx = f0.w * inv.x, sy = f1.x * inv.y, sz = f1.y * inv.z;
    const float bx = fmaf(f0.x, inv.x, -oi.x), by = fmaf(f0.y, inv.y, -oi.y),
                bz = fmaf(f0.z, inv.z, -oi.z);
    const int chs[4] = {int(__float_as_uint(f1.z)), int(__float_as_uint(f1.w)), int(f2.x),
                        int(f2.y)};
    // The decode is monotone in the code (fma with a step of the sign of
    // 1/d), so the entry plane of each axis is the lo code for a positive
    // direction and the hi code for a negative one: select the code words
    // once per node instead of ordering every slab pair.
    const bool nx = inv.x < 0.f, ny = inv.y < 0.f, nz = inv.z < 0.f;
    const uint32_t in_x = nx ? qa.y : qa.x, out_x = nx ? qa.x : qa.y;
    const uint32_t in_y = ny ? qa.w : qa.z, out_y = ny ? qa.z : qa.w;
    const uint32_t in_z = nz ? qb.y : qb.x, out_z = nz ? qb.x : qb.y;
    float dist[4];
    int ref[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t sh = 8u * uint32_t(k);
      const float x0 = fmaf(float((in_x >> sh) & 0xffu), sx, bx);
      const float x1 = fmaf(float((out_x >> sh) & 0xffu), sx, bx);
      const float y0 = fmaf(float((in_y >> sh) & 0xffu), sy, by);
      const float y1 = fmaf(float((out_y >> sh) & 0xffu), sy, by);
      const float z0 = fmaf(float((in_z >> sh) & 0xffu), sz, bz);
      const float z1 = fmaf(float((out_z >> sh) & 0xffu), sz, bz);
      const float tmin = fmaxf(fmaxf(x0, y0), fmaxf(z0, 0.0f));
      const float tmax = fminf(fminf(x1, y1), fminf(z1, best_t));
      const bool hit = tmin <= tmax && chs[k] != kEmptyChildRef;
      dist[k] = hit ? tmin : __builtin_inff();
      ref[k] = chs[k];
    }
    // sort the four slots by entry distance (misses last)
#define PT_CSWAP(i, j)                                                         \
  if (dist[j] < dist[i]) {                                                     \
    const float td = dist[i];                                                  \
    dist[i] = dist[j];                                                         \
    dist[j] = td;                                                              \
    const int tr = ref[i];                                                     \
    ref[i] = ref[j];                                                           \
    ref[j] = tr;                                                               \
  }
    PT_CSWAP(0, 1)
    PT_CSWAP(2, 3)
    PT_CSWAP(0, 2)
    PT_CSWAP(1, 3)
    PT_CSWAP(1, 2)
#undef PT_CSWAP
    if (dist[0] == __builtin_inff()) {
      node = stack.empty() ? kNoRef : stack.pop();
      return;
    }
    if (dist[3] != __builtin_inff())
      stack.push(ref[3]);
    if (dist[2] != __builtin_inff())
      stack.push(ref[2]);
    if (dist[1] != __builtin_inff())
      stack.push(ref[1]);
    node = ref[0];
  }

  // Test every triangle of the leaf `leaf` (= ~first triangle).
  __device__ __forceinline__ void intersect_leaf(const float4 *__restrict__ tris, int leaf) {
    uint32_t ti = uint32_t(~leaf);
    for (;;) {
      const float4 *tp = tris + 3 * ti;
      const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
      const uint32_t word = __float_as_uint(p0.w);
      const uint32_t idx = word & ~kLeafEnd;
      const TriGeom<float> tri = {{p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, {p2.x, p2.y, p2.z}};
      float t;
      uint32_t back;
      if (intersect_triangle_screened(ray, tri, best_t, &t, &back) &&
          hit_before(t, idx, best_t, best_idx)) {
        best_t = t;
        best_idx = idx;
        best_back = back;
      }
      if (word & kLeafEnd)
        break;
      ++ti;
    }
  }

  // One step of a wave whose lanes vote on the kind of step: the wave runs
  // inner-node visits when vote * (lanes at an inner node) >= 4 * (lanes at a
  // leaf), else leaf tests; lanes at the other kind wait. Each wave iteration
  // then executes one of the two code paths instead of both when its lanes
  // are mixed. `has`: this lane holds an unfinished ray (node != kNoRef).
  template <class Stack>
  __device__ __forceinline__ bool step_voted(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ tris, Stack &stack,
                                             bool has, uint32_t vote) {
    const bool at_inner = has && node >= 0;
    const bool at_leaf = has && node < 0;
    const uint32_t ni = uint32_t(__popcll(__ballot(at_inner)));
    const uint32_t nl = uint32_t(__popcll(__ballot(at_leaf)));
    if (vote * ni >= 4u * nl) {
      if (!at_inner)
        return false;
      visit_inner(nodes, stack);
      return node == kNoRef;
    }
    if (!at_leaf)
      return false;
    intersect_leaf(tris, node);
    if (best_idx != kNoHit && best_t < occlude_below)
      return true;
    node = stack.empty() ? kNoRef : stack.pop();
    return node == kNoRef;
  }

  // One step. Returns true when the ray is finished.
  //  kHoldLeaves = false: visit one inner node or one leaf.
  //  kHoldLeaves = true: visit inner nodes until every active lane of the
  //  wave holds a leaf (a lane that already holds one keeps descending), then
  //  test the held leaves together, so the triangle tests run on full waves.
  template <bool kHoldLeaves, class Stack>
  __device__ __forceinline__ bool step(const float4 *__restrict__ nodes,
                                       const float4 *__restrict__ tris, Stack &stack) {
    if constexpr (kHoldLeaves) {
      for (;;) {
        if (pending == kNoRef && node < 0 && node != kNoRef) {
          pending = node;
          node = stack.empty() ? kNoRef : stack.pop();
        }
        if (__ballot(pending == kNoRef && node >= 0) == 0)
          break;
        if (node >= 0)
          visit_inner(nodes, stack);
      }
      if (pending != kNoRef) {
        intersect_leaf(tris, pending);
        pending = kNoRef;
        if (best_idx != kNoHit && best_t < occlude_below)
          return true;
      }
      return node == kNoRef;
    } else {
      if (node >= 0) {
        visit_inner(nodes, stack);
        return node == kNoRef;
      }
      intersect_leaf(tris, node);
      if (best_idx != kNoHit && best_t < occlude_below)
        return true;
      node = stack.empty() ? kNoRef : stack.pop();
      return node == kNoRef;
    }
  }
};

// Whole traversal of one ray, holding leaves until the wave has one each
// (mega-kernel, 82k-triangle scene: 188 vs 175 Msamples/s; 36-triangle
// Cornell box: 688 vs 587). Voting on the step kind per wave iteration
// (step_voted, the wavefront traversal's schedule) measured slower here: Cornell
// box 1,919-1,931 vs 2,174 Msamples/s, 82k-triangle scene 265-270 vs 278.
template <class Stack>
__device__ __forceinline__ Hit<float>
trace_closest(const float4 *__restrict__ nodes, const float4 *__restrict__ tris,
              const Ray<float> &ray, float t_limit, float occlude_below, Stack &stack) {
  TraceState ts;
  ts.init(ray, t_limit, occlude_below);
  stack.sp = 0;
  while (!ts.template step<true>(nodes, tris, stack)) {
  }
  return ts.hit();
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
