#include "bvh.h"

#include <algorithm>
#include <stdexcept>
#include <cmath>
#include <cstring>
#include <limits>

namespace pt {
namespace host {
namespace {

struct Box {
  float lo[3], hi[3];
  static Box empty() {
    Box b;
    for (int i = 0; i < 3; ++i) {
      b.lo[i] = std::numeric_limits<float>::max();
      b.hi[i] = std::numeric_limits<float>::lowest();
    }
    return b;
  }
  void grow(const Box &o) {
    for (int i = 0; i < 3; ++i) {
      lo[i] = std::min(lo[i], o.lo[i]);
      hi[i] = std::max(hi[i], o.hi[i]);
    }
  }
  // reference: include/intersect/accel/aabb.h:80-90 (0 for an inverted box)
  float surface_area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.f || dy < 0.f || dz < 0.f)
      return 0.f;
    return 2 * (dx * dy + dz * dy + dz * dx);
  }
  float centroid(int axis) const { return (lo[axis] + hi[axis]) / 2.f; }
};

struct Builder {
  const std::vector<Box> *bounds;
  BvhSettings settings;
  std::vector<BvhNode> nodes;
  uint32_t max_depth = 0, max_leaf = 0;

  struct Split {
    float cost;
    uint32_t point;
    int axis;
  };

  Split best_split(std::vector<uint32_t> &idx, int axis,
                   std::vector<uint32_t> &sorted_out) {
    sorted_out = idx;
    const auto &bd = *bounds;
    std::stable_sort(sorted_out.begin(), sorted_out.end(),
                     [&](uint32_t l, uint32_t r) {
                       return bd[l].centroid(axis) < bd[r].centroid(axis);
                     });
    size_t n = sorted_out.size();
    std::vector<float> back_sa(n);
    Box run = Box::empty();
    for (size_t i = n; i-- > 0;) {
      run.grow(bd[sorted_out[i]]);
      back_sa[i] = run.surface_area();
    }
    float best = std::numeric_limits<float>::max();
    uint32_t best_i = 0;
    Box left = Box::empty();
    for (size_t i = 0; i < n; ++i) {
      float cost = float(i) * left.surface_area() + float(n - i) * back_sa[i];
      if (cost < best) {
        best = cost;
        best_i = uint32_t(i);
      }
      left.grow(bd[sorted_out[i]]);
    }
    return {best, best_i, axis};
  }

  // Builds node `self` from the triangle ids in idx (reordered in place into
  // leaf order); leaf ranges start at start_idx.
  void create(uint32_t self, std::vector<uint32_t> &idx, uint32_t start_idx,
              uint32_t depth) {
    max_depth = std::max(max_depth, depth);
    const auto &bd = *bounds;
    Box overall = Box::empty();
    for (uint32_t i : idx)
      overall.grow(bd[i]);
    Split best{std::numeric_limits<float>::max(), 0, 0};
    std::vector<uint32_t> best_order, tmp;
    for (int axis = 0; axis < 3; ++axis) {
      Split s = best_split(idx, axis, tmp);
      if (s.cost < best.cost || axis == 0) {
        best = s;
        best_order.swap(tmp);
      }
    }
    float cost = best.cost / overall.surface_area() +
                 settings.traversal_per_intersect_cost;
    BvhNode node;
    node.lo = {overall.lo[0], overall.lo[1], overall.lo[2]};
    node.hi = {overall.hi[0], overall.hi[1], overall.hi[2]};
    if (!(cost < float(idx.size()))) {
      node.is_leaf = 1;
      node.a = start_idx;
      node.b = start_idx + uint32_t(idx.size());
      max_leaf = std::max(max_leaf, uint32_t(idx.size()));
      nodes[self] = node;
      return;
    }
    uint32_t left = uint32_t(nodes.size());
    nodes.resize(nodes.size() + 2);
    uint32_t right = left + 1;
    std::vector<uint32_t> l(best_order.begin(), best_order.begin() + best.point);
    std::vector<uint32_t> r(best_order.begin() + best.point, best_order.end());
    create(left, l, start_idx, depth + 1);
    create(right, r, start_idx + uint32_t(l.size()), depth + 1);
    std::copy(l.begin(), l.end(), idx.begin());
    std::copy(r.begin(), r.end(), idx.begin() + l.size());
    node.is_leaf = 0;
    node.a = left;
    node.b = right;
    nodes[self] = node;
  }
};

} // namespace

Bvh build_sah_bvh(const std::vector<TriGeom<float>> &tris,
                  const BvhSettings &settings) {
  Bvh out;
  if (tris.empty())
    return out;
  std::vector<Box> bounds(tris.size());
  for (size_t i = 0; i < tris.size(); ++i) {
    const auto &t = tris[i];
    V3<float> v[3] = {t.v0, t.v0 + t.e1, t.v0 + t.e2};
    Box b = Box::empty();
    for (auto &p : v) {
      float c[3] = {p.x, p.y, p.z};
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(b.lo[k], c[k]);
        b.hi[k] = std::max(b.hi[k], c[k]);
      }
    }
    bounds[i] = b;
  }
  Builder b;
  b.bounds = &bounds;
  b.settings = settings;
  b.nodes.resize(1);
  std::vector<uint32_t> idx(tris.size());
  for (size_t i = 0; i < idx.size(); ++i)
    idx[i] = uint32_t(i);
  b.create(0, idx, 0, 0);
  out.nodes = std::move(b.nodes);
  out.perm = std::move(idx);
  out.max_depth = b.max_depth;
  out.max_leaf = b.max_leaf;
  return out;
}

float bvh_padding(const std::vector<TriGeom<float>> &tris) {
  float m = 1e-3f;
  for (const auto &t : tris) {
    V3<float> v[3] = {t.v0, t.v0 + t.e1, t.v0 + t.e2};
    for (auto &p : v)
      m = std::max(m, std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z))));
  }
  return m * 0x1p-13f;
}

GpuBvh pack_gpu_bvh(const Bvh &bvh, const std::vector<TriGeom<float>> &tris) {
  GpuBvh g;
  g.pad = bvh_padding(tris);
  g.tris.resize(bvh.perm.size());
  for (size_t i = 0; i < bvh.perm.size(); ++i) {
    const auto &t = tris[bvh.perm[i]];
    GpuTri &o = g.tris[i];
    o.v0[0] = t.v0.x;
    o.v0[1] = t.v0.y;
    o.v0[2] = t.v0.z;
    o.idx = bvh.perm[i];
    o.e1[0] = t.e1.x;
    o.e1[1] = t.e1.y;
    o.e1[2] = t.e1.z;
    o.e2[0] = t.e2.x;
    o.e2[1] = t.e2.y;
    o.e2[2] = t.e2.z;
    o.pad0 = o.pad1 = 0;
  }
  // Mark the last triangle of every leaf.
  for (const BvhNode &n : bvh.nodes)
    if (n.is_leaf && n.b > n.a)
      g.tris[n.b - 1].idx |= 0x80000000u;
  if (bvh.nodes.empty()) {
    // No triangles: one root with four empty slots, so every traversal visits
    // it, hits nothing and ends, without a special case in the kernels.
    GpuNode o{};
    for (int a = 0; a < 3; ++a)
      o.step[a] = 1.0f;
    for (int k = 0; k < 4; ++k) {
      o.child[k] = kEmptyChild;
      o.lo_x[k] = o.lo_y[k] = o.lo_z[k] = 255;
      o.hi_x[k] = o.hi_y[k] = o.hi_z[k] = 0;
    }
    g.nodes.push_back(o);
    g.stack_need = 3;
    return g;
  }
  const float pad = g.pad;
  auto area = [&](uint32_t i) {
    const BvhNode &n = bvh.nodes[i];
    const float dx = n.hi.x - n.lo.x, dy = n.hi.y - n.lo.y, dz = n.hi.z - n.lo.z;
    return dx * dy + dy * dz + dz * dx;
  };
  // Depth-first construction: a GPU node is appended when first reached, so
  // children follow their parents in memory.
  struct Item {
    uint32_t bvh2; // binary node (inner, or a leaf root)
    int32_t slot;  // GPU node index to create
    uint32_t depth;
  };
  std::vector<Item> todo;
  g.nodes.emplace_back();
  todo.push_back({0u, 0, 0u});
  while (!todo.empty()) {
    const Item it = todo.back();
    todo.pop_back();
    // up to four descendants, opening the largest inner child first
    uint32_t kids[4];
    uint32_t nk = 0;
    if (bvh.nodes[it.bvh2].is_leaf) {
      kids[nk++] = it.bvh2;
    } else {
      kids[nk++] = bvh.nodes[it.bvh2].a;
      kids[nk++] = bvh.nodes[it.bvh2].b;
      while (nk < 4) {
        int best = -1;
        float best_area = -1.f;
        for (uint32_t k = 0; k < nk; ++k)
          if (!bvh.nodes[kids[k]].is_leaf && area(kids[k]) > best_area) {
            best = int(k);
            best_area = area(kids[k]);
          }
        if (best < 0)
          break;
        const BvhNode &open = bvh.nodes[kids[best]];
        kids[best] = open.a;
        kids[nk++] = open.b;
      }
    }
    // child boxes (padded), then quantised to the node's grid
    float clo[4][3], chi[4][3];
    int32_t child[4];
    for (uint32_t k = 0; k < nk; ++k) {
      const BvhNode &n = bvh.nodes[kids[k]];
      const float lo[3] = {n.lo.x - pad, n.lo.y - pad, n.lo.z - pad};
      const float hi[3] = {n.hi.x + pad, n.hi.y + pad, n.hi.z + pad};
      for (int a = 0; a < 3; ++a) {
        clo[k][a] = lo[a];
        chi[k][a] = hi[a];
      }
      if (n.is_leaf) {
        child[k] = ~int32_t(n.a);
      } else {
        const int32_t slot = int32_t(g.nodes.size());
        g.nodes.emplace_back();
        child[k] = slot;
        todo.push_back({kids[k], slot, it.depth + 1});
      }
    }
    GpuNode o;
    if (!quantise_gpu_node(clo, chi, nk, child, &o))
      throw std::runtime_error("BVH node quantisation is not conservative");
    g.nodes[size_t(it.slot)] = o;
    g.max_depth = std::max(g.max_depth, it.depth);
  }
  // at most three siblings are pushed per level on the way down
  g.stack_need = 3 * (g.max_depth + 1);
  return g;
}

} // namespace host
} // namespace pt
