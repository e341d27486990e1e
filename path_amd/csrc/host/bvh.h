// Host SAH BVH builder and its repack into the GPU node layout.
#pragma once

#include <cstddef>

#include "../core/bvh_layout.h"
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

// GPU layout (GpuNode, GpuTri, quantise_gpu_node): core/bvh_layout.h.
using ::pt::GpuNode;
using ::pt::GpuTri;
using ::pt::kEmptyChild;

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
