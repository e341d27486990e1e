// GPU BVH record layout and the conservative 8-bit quantisation of a node's
// child boxes, shared by the host packer (host/bvh.cpp) and the GPU builder
// (gpu/lbvh.hip) so both produce the same nodes for the same boxes.
#pragma once

#include "pt_defs.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>

namespace pt {

// GPU layout: a 4-wide BVH collapsed from the binary tree (each GPU node takes
// the up-to-4 descendants obtained by repeatedly opening the child with the
// largest surface area), 64 bytes per node with the four child boxes
// quantised to 8 bits on a per-node grid:
//   origin (x, y, z) and step (sx, sy, sz) floats, each step a power of two
//     -> plane = fma(q, step, origin) per axis,
//   child[4] (>= 0 inner node, < 0 ~(first triangle) of a leaf, kEmptyChild
//     for an unused slot),
//   q lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4] (uint8).
// The grid codes round outward and are checked with the same float fma the
// GPU evaluates, so each decoded box contains its child's padded box: boxes
// only grow, and since the padded boxes are already conservative with
// respect to the float triangle test, the closest hit still does not depend
// on the tree.
constexpr int32_t kEmptyChild = -2147483647 - 1; // INT32_MIN

struct GpuNode {
  float origin[3];
  float step[3];
  int32_t child[4];
  uint8_t lo_x[4], hi_x[4], lo_y[4], hi_y[4];
  uint8_t lo_z[4], hi_z[4];
};
static_assert(sizeof(GpuNode) == 64, "GPU BVH node is one 64-byte record");
// gpu/trace.h reads the node as four float4: origin + step x, steps y z +
// child 0 1, child 2 3 + lo_x hi_x, lo_y hi_y lo_z hi_z
static_assert(offsetof(GpuNode, step) == 12 && offsetof(GpuNode, child) == 24 &&
                  offsetof(GpuNode, lo_x) == 40 && offsetof(GpuNode, lo_y) == 48 &&
                  offsetof(GpuNode, lo_z) == 56 && offsetof(GpuNode, hi_z) == 60,
              "GpuNode layout read by the traversal");

// Triangle in leaf order: (v0, idx bits), (e1, 0), (e2, 0).
struct GpuTri {
  float v0[3];
  uint32_t idx;
  float e1[3];
  uint32_t pad0;
  float e2[3];
  uint32_t pad1;
};

// Quantise up to four child boxes (already padded) onto the node's grid:
// origin = the union's low corner, per axis the smallest power-of-two step
// with 254 steps covering the extent (one step of slack so outward rounding
// never needs code 256), codes rounded outward and checked with the float fma
// the traversal decodes with. Empty slots get lo = 255, hi = 0. Returns false
// if a decoded box would not contain its child's box.
PT_INLINE bool quantise_gpu_node(const float (*clo)[3], const float (*chi)[3], uint32_t nk,
                                 const int32_t *child, GpuNode *out) {
  float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t k = 0; k < nk; ++k)
    for (int a = 0; a < 3; ++a) {
      nlo[a] = fminf(nlo[a], clo[k][a]);
      nhi[a] = fmaxf(nhi[a], chi[k][a]);
    }
  GpuNode o{};
  uint8_t *qlo[3] = {o.lo_x, o.lo_y, o.lo_z}, *qhi[3] = {o.hi_x, o.hi_y, o.hi_z};
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const float extent = nhi[a] - nlo[a];
    int e = -126;
    if (extent > 0.f) {
      int ex;
      frexpf(extent / 254.0f, &ex);
      e = ex > -126 ? ex : -126;
      while (ldexpf(254.0f, e - 1) >= extent && e > -126)
        --e;
    }
    const float scale = ldexpf(1.0f, e);
    o.origin[a] = nlo[a];
    o.step[a] = scale;
    for (uint32_t k = 0; k < 4; ++k) {
      if (k >= nk) {
        qlo[a][k] = 255;
        qhi[a][k] = 0;
        continue;
      }
      int ql = int(floorf((clo[k][a] - nlo[a]) / scale));
      int qh = int(ceilf((chi[k][a] - nlo[a]) / scale));
      ql = ql < 0 ? 0 : (ql > 255 ? 255 : ql);
      qh = qh < 0 ? 0 : (qh > 255 ? 255 : qh);
      // outward on the exact float decode
      while (ql > 0 && fmaf(float(ql), scale, nlo[a]) > clo[k][a])
        --ql;
      while (qh < 255 && fmaf(float(qh), scale, nlo[a]) < chi[k][a])
        ++qh;
      ok = ok && !(fmaf(float(ql), scale, nlo[a]) > clo[k][a] ||
                   fmaf(float(qh), scale, nlo[a]) < chi[k][a]);
      qlo[a][k] = uint8_t(ql);
      qhi[a][k] = uint8_t(qh);
    }
  }
  for (uint32_t k = 0; k < 4; ++k)
    o.child[k] = k < nk ? child[k] : kEmptyChild;
  *out = o;
  return ok;
}

} // namespace pt
