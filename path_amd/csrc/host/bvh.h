// Host SAH BVH builder and its repack into the GPU node layout.
#pragma once

#include <cstddef>

#include "../core/geometry.h"

#include <climits>
#include <cstdint>
#include <vector>

namespace pt {
namespace host {

// Binary BVH in the reference's layout (include/intersect/accel/detail/bvh/node.h:
// 22-126): node 0 is the root, children are allocated in pairs, a leaf covers
// the contiguous range [start, end) of the permuted triangle array.
struct BvhNode {
  V3<float> lo, hi;
  uint32_t is_leaf;
  uint32_t a; // inner: left child; leaf: start
  uint32_t b; // inner: right child; leaf: end
};

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> perm; // perm[i] = flattened index of the i-th leaf-ordered triangle
  uint32_t max_depth = 0;
  uint32_t max_leaf = 0;
};

struct BvhSettings {
  float traversal_per_intersect_cost = 0.5f; // reference default (sbvh/settings.h)
};

// Traversal cost of the tree the GPU traverses. Measured on MI355X (82k-
// triangle scene, quantised 4-wide nodes, screened triangle test): 0.25 ->
// 487, 0.375 -> 490, 0.5 -> 488, 0.75 -> 484, 1.0 -> 475, 1.5 -> 376
// Msamples/s, so the reference default is kept (with 128-byte float nodes
// and the full triangle test, 1.0 had been best). Closest hits do not depend
// on the tree, so images are unchanged.
constexpr float kGpuSahTraversalCost = 0.5f;

// SAH object-split builder following src/intersect/accel/sbvh/detail/generator.cu:
// 16-438 (spatial splits off, the reference default): per node and axis, sort
// by bound centroid, sweep all split points with cost i*SA_L + (n-i)*SA_R, make
// a leaf when cost/SA + traversal_cost >= n. Centroid ties keep input order
// (the reference's std::sort leaves them unspecified).
Bvh build_sah_bvh(const std::vector<TriGeom<float>> &tris,
                  const BvhSettings &settings = {});

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
constexpr int32_t kEmptyChild = INT32_MIN;

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

struct GpuBvh {
  std::vector<GpuNode> nodes;
  std::vector<GpuTri> tris;
  float pad = 0.f;
  uint32_t max_depth = 0;   // depth of the 4-wide tree (root = 0)
  uint32_t stack_need = 0;  // traversal stack entries a ray can need
};

GpuBvh pack_gpu_bvh(const Bvh &bvh, const std::vector<TriGeom<float>> &tris);

// Conservative padding for a scene: a relative 2^-13 of its largest
// coordinate magnitude.
float bvh_padding(const std::vector<TriGeom<float>> &tris);

} // namespace host
} // namespace pt
