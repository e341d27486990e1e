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

// One reference to a triangle: its id and the bounds of the part of it this
// reference stands for (the whole triangle until a spatial split cuts it).
struct Ref {
  Box box;
  uint32_t tri;
};

Box intersect(const Box &a, const Box &b) {
  Box o;
  for (int i = 0; i < 3; ++i) {
    o.lo[i] = std::max(a.lo[i], b.lo[i]);
    o.hi[i] = std::min(a.hi[i], b.hi[i]);
  }
  return o;
}

bool is_empty(const Box &b) {
  return b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2];
}

// Spatial-split search resolution: planes per axis of a node with n
// references (2n, as the reference's generator.cu:443, capped so a node's
// search stays linear in n).
constexpr uint32_t kMaxSpatialBins = 256;
// No spatial splits below this depth of the binary tree: the GPU's traversal
// stack bounds the collapsed tree's depth.
constexpr uint32_t kMaxSpatialDepth = 40;

struct Builder {
  const std::vector<TriGeom<float>> *tris;
  BvhSettings settings;
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> perm;
  uint32_t max_depth = 0, max_leaf = 0, n_spatial = 0;
  float root_sa = 0.f;
  size_t n_refs = 0, ref_budget = 0;

  struct Split {
    float cost;
    uint32_t point;
    int axis;
    float overlap_sa; // surface area of left box ∩ right box
  };

  Split best_object_split(const std::vector<Ref> &refs, int axis,
                          std::vector<uint32_t> &sorted_out) {
    sorted_out.resize(refs.size());
    for (size_t i = 0; i < refs.size(); ++i)
      sorted_out[i] = uint32_t(i);
    std::stable_sort(sorted_out.begin(), sorted_out.end(), [&](uint32_t l, uint32_t r) {
      return refs[l].box.centroid(axis) < refs[r].box.centroid(axis);
    });
    size_t n = sorted_out.size();
    std::vector<Box> back(n);
    Box run = Box::empty();
    for (size_t i = n; i-- > 0;) {
      run.grow(refs[sorted_out[i]].box);
      back[i] = run;
    }
    float best = std::numeric_limits<float>::max();
    uint32_t best_i = 0;
    float best_overlap = 0.f;
    Box left = Box::empty();
    for (size_t i = 0; i < n; ++i) {
      float cost = float(i) * left.surface_area() + float(n - i) * back[i].surface_area();
      if (cost < best) {
        best = cost;
        best_i = uint32_t(i);
        best_overlap = intersect(left, back[i]).surface_area();
      }
      left.grow(refs[sorted_out[i]].box);
    }
    return {best, best_i, axis, best_overlap};
  }

  // Part of reference r between axis planes a <= b.
  Box chop(const Ref &r, int axis, float a, float b) const {
    Box slab = r.box;
    slab.lo[axis] = std::max(slab.lo[axis], a);
    slab.hi[axis] = std::min(slab.hi[axis], b);
    const TriGeom<float> &t = (*tris)[r.tri];
    const V3<float> v[3] = {t.v0, t.v0 + t.e1, t.v0 + t.e2};
    Box o;
    clip_triangle_bounds(v, slab.lo, slab.hi, o.lo, o.hi);
    // Rounding can leave a sliver empty; its slab bound is conservative.
    return is_empty(o) ? slab : o;
  }

  struct Spatial {
    float cost = std::numeric_limits<float>::max();
    int axis = 0;
    float plane = 0.f;
    std::vector<Ref> left, right;
  };

  // Binned spatial split on one axis (Stich, Friedrich, Dietrich, "Spatial
  // Splits in Bounding Volume Hierarchies", HPG 2009): references are chopped
  // into the bins they cross; a plane's cost counts a reference on each side
  // it enters. The chosen plane's straddlers are then offered "unsplitting"
  // (kept whole on one side when that is cheaper), and the cost returned is
  // that of the final partition.
  void spatial_split(const std::vector<Ref> &refs, const Box &overall, int axis,
                     Spatial &best) const {
    const float lo = overall.lo[axis], hi = overall.hi[axis];
    if (!(hi > lo))
      return;
    const uint32_t nb = uint32_t(std::min<size_t>(kMaxSpatialBins, 2 * refs.size()));
    std::vector<float> plane(nb + 1);
    for (uint32_t k = 0; k < nb; ++k)
      plane[k] = lo + (hi - lo) * (float(k) / float(nb));
    plane[nb] = hi;
    // lower bin: plane[b] <= x; upper bin: x <= plane[b + 1]
    auto bin_lo = [&](float x) {
      int b = int(std::floor((x - lo) / (hi - lo) * float(nb)));
      b = std::min(std::max(b, 0), int(nb) - 1);
      while (b > 0 && plane[b] > x)
        --b;
      while (b + 1 < int(nb) && plane[b + 1] <= x)
        ++b;
      return b;
    };
    auto bin_hi = [&](float x) {
      int b = int(std::floor((x - lo) / (hi - lo) * float(nb)));
      b = std::min(std::max(b, 0), int(nb) - 1);
      while (b + 1 < int(nb) && plane[b + 1] < x)
        ++b;
      while (b > 0 && plane[b] >= x)
        --b;
      return b;
    };
    std::vector<Box> bins(nb, Box::empty());
    std::vector<uint32_t> enter(nb, 0), leave(nb, 0);
    for (const Ref &r : refs) {
      const int b0 = bin_lo(r.box.lo[axis]);
      const int b1 = std::max(b0, bin_hi(r.box.hi[axis]));
      ++enter[b0];
      ++leave[b1];
      if (b0 == b1) {
        bins[b0].grow(r.box);
        continue;
      }
      for (int k = b0; k <= b1; ++k)
        bins[k].grow(chop(r, axis, plane[k], plane[k + 1]));
    }
    std::vector<Box> back(nb);
    Box run = Box::empty();
    for (uint32_t k = nb; k-- > 0;) {
      run.grow(bins[k]);
      back[k] = run;
    }
    Box left = Box::empty();
    uint32_t n_left = 0, n_right = uint32_t(refs.size());
    float est = std::numeric_limits<float>::max();
    uint32_t best_k = 0;
    for (uint32_t k = 1; k < nb; ++k) {
      left.grow(bins[k - 1]);
      n_left += enter[k - 1];
      n_right -= leave[k - 1];
      const float cost = float(n_left) * left.surface_area() +
                         float(n_right) * back[k].surface_area();
      if (cost < est) {
        est = cost;
        best_k = k;
      }
    }
    if (best_k == 0 || !(est < best.cost))
      return;
    // partition at plane[best_k]
    const float p = plane[best_k];
    Spatial s;
    s.axis = axis;
    s.plane = p;
    Box bl = Box::empty(), br = Box::empty();
    std::vector<uint32_t> straddle;
    for (uint32_t i = 0; i < refs.size(); ++i) {
      const Ref &r = refs[i];
      if (r.box.lo[axis] >= p) {
        s.right.push_back(r);
        br.grow(r.box);
      } else if (r.box.hi[axis] <= p) {
        s.left.push_back(r);
        bl.grow(r.box);
      } else {
        straddle.push_back(i);
      }
    }
    std::vector<Box> cut_l(straddle.size()), cut_r(straddle.size());
    for (size_t j = 0; j < straddle.size(); ++j) {
      const Ref &r = refs[straddle[j]];
      cut_l[j] = chop(r, axis, std::numeric_limits<float>::lowest(), p);
      cut_r[j] = chop(r, axis, p, std::numeric_limits<float>::max());
      bl.grow(cut_l[j]);
      br.grow(cut_r[j]);
    }
    float nl = float(s.left.size() + straddle.size());
    float nr = float(s.right.size() + straddle.size());
    for (size_t j = 0; j < straddle.size(); ++j) {
      const Ref &r = refs[straddle[j]];
      Box ul = bl, ur = br;
      ul.grow(r.box);
      ur.grow(r.box);
      const float c_split = nl * bl.surface_area() + nr * br.surface_area();
      const float c_left = nl * ul.surface_area() + (nr - 1.f) * br.surface_area();
      const float c_right = (nl - 1.f) * bl.surface_area() + nr * ur.surface_area();
      if (c_left < c_split && c_left <= c_right) {
        s.left.push_back(r);
        bl = ul;
        nr -= 1.f;
      } else if (c_right < c_split) {
        s.right.push_back(r);
        br = ur;
        nl -= 1.f;
      } else {
        s.left.push_back({cut_l[j], r.tri});
        s.right.push_back({cut_r[j], r.tri});
      }
    }
    if (s.left.empty() || s.right.empty())
      return;
    Box fl = Box::empty(), fr = Box::empty();
    for (const Ref &r : s.left)
      fl.grow(r.box);
    for (const Ref &r : s.right)
      fr.grow(r.box);
    s.cost = float(s.left.size()) * fl.surface_area() + float(s.right.size()) * fr.surface_area();
    if (s.cost < best.cost)
      best = std::move(s);
  }

  // Builds node `self` from the references in refs; leaves append their
  // triangles to perm in depth-first order, so every leaf covers a
  // contiguous range.
  void create(uint32_t self, std::vector<Ref> &refs, uint32_t depth) {
    max_depth = std::max(max_depth, depth);
    Box overall = Box::empty();
    for (const Ref &r : refs)
      overall.grow(r.box);
    if (depth == 0)
      root_sa = overall.surface_area();
    Split best{std::numeric_limits<float>::max(), 0, 0, 0.f};
    std::vector<uint32_t> best_order, tmp;
    for (int axis = 0; axis < 3; ++axis) {
      Split s = best_object_split(refs, axis, tmp);
      if (s.cost < best.cost || axis == 0) {
        best = s;
        best_order.swap(tmp);
      }
    }
    // reference generator.cu:112-126: spatial splits are tried when the best
    // object split's children overlap by more than alpha x the root's area
    Spatial sp;
    if (settings.spatial_splits && depth < kMaxSpatialDepth && n_refs < ref_budget &&
        best.overlap_sa / root_sa > settings.overlap_threshold) {
      sp.cost = best.cost;
      for (int axis = 0; axis < 3; ++axis)
        spatial_split(refs, overall, axis, sp);
    }
    const bool use_spatial = !sp.left.empty();
    const float split_cost = use_spatial ? sp.cost : best.cost;
    float cost = split_cost / overall.surface_area() + settings.traversal_per_intersect_cost;
    BvhNode node;
    node.lo = {overall.lo[0], overall.lo[1], overall.lo[2]};
    node.hi = {overall.hi[0], overall.hi[1], overall.hi[2]};
    if (!(cost < float(refs.size()))) {
      node.is_leaf = 1;
      node.a = uint32_t(perm.size());
      for (const Ref &r : refs)
        perm.push_back(r.tri);
      node.b = uint32_t(perm.size());
      max_leaf = std::max(max_leaf, uint32_t(refs.size()));
      nodes[self] = node;
      return;
    }
    uint32_t left = uint32_t(nodes.size());
    nodes.resize(nodes.size() + 2);
    uint32_t right = left + 1;
    std::vector<Ref> l, r;
    if (use_spatial) {
      ++n_spatial;
      n_refs += sp.left.size() + sp.right.size() - refs.size();
      l = std::move(sp.left);
      r = std::move(sp.right);
    } else {
      l.reserve(best.point);
      r.reserve(refs.size() - best.point);
      for (uint32_t i = 0; i < best.point; ++i)
        l.push_back(refs[best_order[i]]);
      for (size_t i = best.point; i < refs.size(); ++i)
        r.push_back(refs[best_order[i]]);
    }
    std::vector<Ref>().swap(refs);
    create(left, l, depth + 1);
    std::vector<Ref>().swap(l);
    create(right, r, depth + 1);
    node.is_leaf = 0;
    node.a = left;
    node.b = right;
    nodes[self] = node;
  }
};

} // namespace

void clip_triangle_bounds(const V3<float> v[3], const float lo[3], const float hi[3],
                          float out_lo[3], float out_hi[3]) {
  // At most 3 + 6 vertices after six planes.
  double poly[2][12][3];
  int n = 3, cur = 0;
  for (int k = 0; k < 3; ++k) {
    poly[0][k][0] = v[k].x;
    poly[0][k][1] = v[k].y;
    poly[0][k][2] = v[k].z;
  }
  for (int axis = 0; axis < 3 && n > 0; ++axis)
    for (int side = 0; side < 2 && n > 0; ++side) {
      const double p = side ? hi[axis] : lo[axis];
      auto inside = [&](const double *q) { return side ? q[axis] <= p : q[axis] >= p; };
      const double(*in)[3] = poly[cur];
      double(*out)[3] = poly[cur ^ 1];
      int m = 0;
      for (int i = 0; i < n; ++i) {
        const double *a = in[i], *b = in[(i + 1) % n];
        const bool ia = inside(a), ib = inside(b);
        if (ia) {
          out[m][0] = a[0], out[m][1] = a[1], out[m][2] = a[2];
          ++m;
        }
        if (ia != ib) {
          const double t = (p - a[axis]) / (b[axis] - a[axis]);
          for (int c = 0; c < 3; ++c)
            out[m][c] = a[c] + t * (b[c] - a[c]);
          out[m][axis] = p;
          ++m;
        }
      }
      n = m;
      cur ^= 1;
    }
  for (int c = 0; c < 3; ++c) {
    out_lo[c] = std::numeric_limits<float>::max();
    out_hi[c] = std::numeric_limits<float>::lowest();
  }
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      const float x = float(poly[cur][i][c]);
      out_lo[c] = std::min(out_lo[c], x);
      out_hi[c] = std::max(out_hi[c], x);
    }
  if (n == 0)
    return;
  for (int c = 0; c < 3; ++c) {
    out_lo[c] = std::max(out_lo[c], lo[c]);
    out_hi[c] = std::min(out_hi[c], hi[c]);
  }
}

Bvh build_sah_bvh(const std::vector<TriGeom<float>> &tris,
                  const BvhSettings &settings) {
  Bvh out;
  if (tris.empty())
    return out;
  std::vector<Ref> refs(tris.size());
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
    refs[i] = {b, uint32_t(i)};
  }
  Builder b;
  b.tris = &tris;
  b.settings = settings;
  b.nodes.resize(1);
  b.perm.reserve(tris.size());
  b.n_refs = tris.size();
  // duplicates stop once references reach twice the triangle count
  b.ref_budget = 2 * tris.size();
  b.create(0, refs, 0);
  out.nodes = std::move(b.nodes);
  out.perm = std::move(b.perm);
  out.max_depth = b.max_depth;
  out.max_leaf = b.max_leaf;
  out.spatial_splits = b.n_spatial;
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
