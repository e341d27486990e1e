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
  // perm[i] = flattened index of the i-th leaf-ordered triangle reference
  // (with spatial splits a triangle can be referenced by several leaves)
  std::vector<uint32_t> perm;
  uint32_t max_depth = 0;
  uint32_t max_leaf = 0;
  uint32_t spatial_splits = 0; // nodes split by a spatial split
};

struct BvhSettings {
  float traversal_per_intersect_cost = 0.5f; // reference default (sbvh/settings.h)
  // Spatial splits (SBVH; reference sbvh/settings.h:13-14, off by default):
  // a node also tries splitting planes that cut triangles when the children
  // of its best object split overlap by more than overlap_threshold times the
  // root's surface area. A cut triangle is referenced by both children.
  // Measured on MI355X, 82k-triangle scene, wavefront integrator: object
  // splits 756-758 Msamples/s; spatial splits at overlap thresholds 1e-5 /
  // 1e-4 / 1e-3 / 1e-2 (95,930 / 82,751 / 82,513 / 82,224 references):
  // 743 / 736 / 735 / 734. So they stay off by default, as in the reference.
  bool spatial_splits = false;
  float overlap_threshold = 1e-5f;
  bool operator==(const BvhSettings &o) const {
    return traversal_per_intersect_cost == o.traversal_per_intersect_cost &&
           spatial_splits == o.spatial_splits && overlap_threshold == o.overlap_threshold;
  }
};

// Bounds of the part of triangle v[0..2] inside the box [lo, hi] (inclusive
// planes; empty bounds, lo > hi, if they do not meet): the triangle polygon
// clipped by the six planes, Sutherland-Hodgman. The spatial-split builder
// chops a reference at a plane with it (reference ClippedTriangle::new_bounds,
// include/intersect/accel/detail/clipped_triangle_impl.h:13-124).
void clip_triangle_bounds(const V3<float> v[3], const float lo[3], const float hi[3],
                          float out_lo[3], float out_hi[3]);

// Traversal cost of the tree the GPU traverses. Measured on MI355X (82k-
// triangle scene, quantised 4-wide nodes, screened triangle test): 0.25 ->
// 487, 0.375 -> 490, 0.5 -> 488, 0.75 -> 484, 1.0 -> 475, 1.5 -> 376
// Msamples/s, so the reference default is kept (with 128-byte float nodes
// and the full triangle test, 1.0 had been best). Closest hits do not depend
// on the tree, so images are unchanged.
constexpr float kGpuSahTraversalCost = 0.5f;

// SAH builder following src/intersect/accel/sbvh/detail/generator.cu:16-713:
// per node and axis, sort the references by bound centroid, sweep all object
// split points with cost i*SA_L + (n-i)*SA_R; with spatial splits, also 32
// planes per axis over the node's extent with the references chopped into
// the bins they cross; make a leaf when cost/SA + traversal_cost >= n.
// Centroid ties keep input order (the reference's std::sort leaves them
// unspecified).
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
