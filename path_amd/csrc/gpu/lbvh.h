// GPU BVH build for batched rendering: per scene, triangles sorted by the
// 30-bit Morton code of their box centre within the scene's centroid bounds
// (scene id in the key's high word, one radix sort for the whole batch), then
// a binary tree over the sorted triangles by one of two builders:
//  * LBVH: the binary radix tree of the sorted codes (Karras, HPG 2012: one
//    thread per internal node, ties broken by position), boxes bottom-up (the
//    second child to arrive at a node computes it);
//  * PLOC (Meister and Bittner, "Parallel Locally-Ordered Clustering for
//    Bounding Volume Hierarchy Construction", IEEE TVCG 2018): clusters in
//    Morton order, each pass pairs every cluster with the neighbour within
//    kPlocRadius positions whose union box has the least surface area and
//    merges the mutual pairs, until one cluster is left (one workgroup per
//    scene). Its trees are closer to SAH trees than the radix tree's.
// Then the 4-wide node records top-down, one level per launch: each node
// opens its largest-area inner child until it has four children, subtrees of
// at most kLbvhLeafSize triangles become leaves whose triangles the level
// writes at the node's output range, and the child boxes are padded and
// quantised with the host packer's rule (core/bvh_layout.h). Closest hits do
// not depend on the tree (padded conservative boxes, order-independent hit
// rule), so a batch renders the same images with these trees as with the
// host SAH tree.
//
// Numbering: the batch's triangles are global positions 0..n_tris-1 (scene s
// owns [tri_begin[s], tri_begin[s+1]) in sorted and in output order); scene s
// owns internal nodes [int_begin[s], int_begin[s+1]), n - 1 of them for n >= 2
// triangles, its root first. A binary child ref >= 0 is an internal node,
// < 0 is ~(sorted triangle position).
#pragma once

#include "../core/bvh_layout.h"
#include "../core/shading.h"

namespace pt {
namespace gpu {

// Subtrees of at most this many triangles become leaves. Measured on 1024
// generated scenes (render of a batch, ms): PLOC leaves of 1 / 2 / 3 / 4 / 6
// triangles 153.4 / 146.6 / 148.0 / 156.0 / 162.8; radius 4 / 8 / 16 with
// leaves of 4: 157.4 / 156.0 / 155.3 (16 adds 0.7 ms of build); the radix
// tree with leaves of 2 / 4: 154.5 / 164.6.
constexpr uint32_t kLbvhLeafSize = 2;
constexpr uint32_t kLbvhLeafEnd = 0x80000000u; // last triangle of a leaf (gpu/trace.h)
constexpr int kPlocRadius = 8;
// PLOC runs a scene in one workgroup; a batch with a larger scene is built
// with the radix tree (one thread per node), which spreads a big scene over
// the whole GPU.
constexpr uint32_t kPlocMaxSceneTris = 1u << 18;
// traversal stack entries a 4-wide tree of this many levels can need
constexpr uint32_t kLbvhMaxLevels = 21; // 3 * (20 + 1) = 63 <= 64

struct LbvhArgs {
  const ShadeTri<float> *shade; // all scenes, flattened order per scene
  const uint32_t *tri_begin;    // [n_scenes + 1]
  const uint32_t *int_begin;    // [n_scenes + 1]
  uint32_t n_scenes, n_tris, n_internal;
  float *scene_box;             // [n_scenes][8]: centroid lo xyz, hi xyz, pad, -
  unsigned long long *keys, *keys_sorted;
  uint32_t *vals, *vals_sorted; // global triangle index
  int32_t *left, *right;        // [n_internal] binary child refs
  uint32_t *count;              // [n_internal] triangles under the node
  int32_t *parent_int;          // [n_internal], -1 at a root
  int32_t *parent_leaf;         // [n_tris], -1 for a single-triangle scene
  float *ibox;                  // [n_internal][6] lo xyz, hi xyz
  float *lbox;                  // [n_tris][6]
  uint32_t *visits;             // [n_internal]
  float *clusters[2];           // [n_tris][8] PLOC clusters: lo xyz + ref, hi xyz + count
  int32_t *nn;                  // [n_tris] PLOC nearest neighbours
  GpuNode *nodes;               // output, global slots
  GpuTri *tris;                 // output, sorted positions
  uint32_t *counters;           // [0] node slots handed out, [2] error bits
};

PT_INLINE uint32_t lbvh_find(const uint32_t *begin, uint32_t n, uint32_t x) {
  // largest s < n with begin[s] <= x (begin non-decreasing, begin[0] = 0)
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (begin[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

PT_INLINE uint32_t lbvh_spread(uint32_t v) { // 10 bits -> every third bit
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Box of flattened triangle t (vertices v0, v0 + e1, v0 + e2 as the host).
PT_INLINE void lbvh_tri_box(const ShadeTri<float> &t, float lo[3], float hi[3]) {
  const V3<float> v[3] = {t.v0, t.v0 + t.e1, t.v0 + t.e2};
  lo[0] = hi[0] = v[0].x;
  lo[1] = hi[1] = v[0].y;
  lo[2] = hi[2] = v[0].z;
  for (int k = 1; k < 3; ++k) {
    const float c[3] = {v[k].x, v[k].y, v[k].z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], c[a]);
      hi[a] = fmaxf(hi[a], c[a]);
    }
  }
}

// Morton key of global triangle t (scene in the high word).
PT_INLINE unsigned long long lbvh_key(const LbvhArgs &a, uint32_t t) {
  const uint32_t s = lbvh_find(a.tri_begin, a.n_scenes, t);
  float lo[3], hi[3];
  lbvh_tri_box(a.shade[t], lo, hi);
  const float *sb = a.scene_box + 8 * size_t(s);
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float c = (lo[k] + hi[k]) * 0.5f;
    const float ext = sb[3 + k] - sb[k];
    float x = ext > 0.f ? (c - sb[k]) / ext : 0.f;
    x = fminf(fmaxf(x * 1024.f, 0.f), 1023.f);
    q[k] = uint32_t(x);
  }
  const uint32_t m = (lbvh_spread(q[0]) << 2) | (lbvh_spread(q[1]) << 1) | lbvh_spread(q[2]);
  return ((unsigned long long)s << 32) | m;
}

// Common-prefix length of sorted positions i and j of a scene (n keys from
// `k`), -1 outside [0, n); equal codes compare their positions.
PT_INLINE int lbvh_delta(const unsigned long long *k, int n, int i, int j) {
  if (j < 0 || j >= n)
    return -1;
  const uint32_t a = uint32_t(k[i]), b = uint32_t(k[j]);
  if (a == b)
    return 32 + __builtin_clz(uint32_t(i ^ j));
  return __builtin_clz(a ^ b);
}

// Internal node g (global): children, range and parents.
PT_INLINE void lbvh_karras(const LbvhArgs &a, uint32_t g) {
  const uint32_t s = lbvh_find(a.int_begin, a.n_scenes, g);
  const uint32_t base = a.tri_begin[s];
  const int n = int(a.tri_begin[s + 1] - base);
  const int i = int(g - a.int_begin[s]);
  const unsigned long long *k = a.keys_sorted + base;
  const int d = lbvh_delta(k, n, i, i + 1) - lbvh_delta(k, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = lbvh_delta(k, n, i, i - d);
  int lmax = 2;
  while (lbvh_delta(k, n, i, i + lmax * d) > dmin)
    lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (lbvh_delta(k, n, i, i + (l + t) * d) > dmin)
      l += t;
  const int j = i + l * d;
  const int dnode = lbvh_delta(k, n, i, j);
  int split = 0, t = l;
  do {
    t = (t + 1) / 2;
    if (lbvh_delta(k, n, i, i + (split + t) * d) > dnode)
      split += t;
  } while (t > 1);
  const int gamma = i + split * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t ib = a.int_begin[s];
  int32_t lref, rref;
  if (lo == gamma) {
    lref = ~int32_t(base + uint32_t(gamma));
    a.parent_leaf[base + uint32_t(gamma)] = int32_t(g);
  } else {
    lref = int32_t(ib + uint32_t(gamma));
    a.parent_int[ib + uint32_t(gamma)] = int32_t(g);
  }
  if (hi == gamma + 1) {
    rref = ~int32_t(base + uint32_t(gamma + 1));
    a.parent_leaf[base + uint32_t(gamma + 1)] = int32_t(g);
  } else {
    rref = int32_t(ib + uint32_t(gamma + 1));
    a.parent_int[ib + uint32_t(gamma + 1)] = int32_t(g);
  }
  a.left[g] = lref;
  a.right[g] = rref;
  a.count[g] = uint32_t(hi - lo + 1);
}

} // namespace gpu
} // namespace pt

namespace pt {
namespace gpu {

struct LbvhRoot {
  int32_t ref;        // binary root (internal node, or ~position of a lone triangle)
  uint32_t slot;      // its node slot (= the scene's node offset)
  uint32_t scene;
  uint32_t root_slot; // slot of the scene's root (child refs are relative to it)
  uint32_t tri_first; // output position of the subtree's first triangle
};

// Box of the triangle at sorted position p.
PT_INLINE void lbvh_leaf_box(const LbvhArgs &a, uint32_t p) {
  float lo[3], hi[3];
  lbvh_tri_box(a.shade[a.vals_sorted[p]], lo, hi);
  float *lb = a.lbox + 6 * size_t(p);
  for (int k = 0; k < 3; ++k) {
    lb[k] = lo[k];
    lb[3 + k] = hi[k];
  }
}

// Sorted position p (LBVH): the leaf box, then the boxes of the ancestors
// this thread is the second to reach. Ops: fence(), add(u32*, v) (returns the
// old value), or(u32*, v).
template <class Ops>
PT_INLINE void lbvh_leaf(const LbvhArgs &a, uint32_t p, Ops &ops) {
  lbvh_leaf_box(a, p);
  int32_t node = a.parent_leaf[p];
  while (node >= 0) {
    ops.fence(); // publish this child's box before arriving
    if (ops.add(&a.visits[node], 1u) == 0u)
      return; // the sibling subtree is not done yet: its last thread goes on
    ops.fence(); // see the sibling's box
    const int32_t c[2] = {a.left[node], a.right[node]};
    float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int j = 0; j < 2; ++j) {
      const float *cb = c[j] < 0 ? a.lbox + 6 * size_t(~c[j]) : a.ibox + 6 * size_t(c[j]);
      for (int k = 0; k < 3; ++k) {
        blo[k] = fminf(blo[k], cb[k]);
        bhi[k] = fmaxf(bhi[k], cb[3 + k]);
      }
    }
    float *ib = a.ibox + 6 * size_t(node);
    for (int k = 0; k < 3; ++k) {
      ib[k] = blo[k];
      ib[3 + k] = bhi[k];
    }
    node = a.parent_int[node];
  }
}

PT_INLINE uint32_t lbvh_count(const LbvhArgs &a, int32_t ref) {
  return ref < 0 ? 1u : a.count[ref];
}

PT_INLINE bool lbvh_leafish(const LbvhArgs &a, int32_t ref) {
  return lbvh_count(a, ref) <= kLbvhLeafSize;
}

// Write the triangles under `ref` (at most kLbvhLeafSize) as one leaf at
// output positions out, out + 1, ...; the last one carries the leaf-end bit.
PT_INLINE void lbvh_emit_leaf(const LbvhArgs &a, int32_t ref, uint32_t out, uint32_t scene) {
  int32_t stack[kLbvhLeafSize];
  uint32_t sp = 0, q = out;
  const uint32_t n = lbvh_count(a, ref), t_base = a.tri_begin[scene];
  stack[sp++] = ref;
  while (sp > 0) {
    const int32_t r = stack[--sp];
    if (r >= 0) { // at most n - 1 inner nodes, each pushing two refs
      if (sp + 2 > kLbvhLeafSize)
        return; // unreachable for a subtree of at most kLbvhLeafSize leaves
      stack[sp++] = a.right[r];
      stack[sp++] = a.left[r];
      continue;
    }
    const uint32_t t = a.vals_sorted[uint32_t(~r)];
    const ShadeTri<float> &st = a.shade[t];
    GpuTri o;
    o.v0[0] = st.v0.x, o.v0[1] = st.v0.y, o.v0[2] = st.v0.z;
    o.idx = (t - t_base) | (q + 1 == out + n ? kLbvhLeafEnd : 0u);
    o.e1[0] = st.e1.x, o.e1[1] = st.e1.y, o.e1[2] = st.e1.z;
    o.e2[0] = st.e2.x, o.e2[1] = st.e2.y, o.e2[2] = st.e2.z;
    o.pad0 = o.pad1 = 0;
    a.tris[q++] = o;
  }
}

PT_INLINE float lbvh_area(const LbvhArgs &a, int32_t ref) {
  const float *b = a.ibox + 6 * size_t(ref);
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

// The up-to-four descendants a level item's node takes (largest-area inner
// child opened first); returns how many of them are inner (not leaves).
PT_INLINE uint32_t lbvh_level_kids(const LbvhArgs &a, const LbvhRoot &it, int32_t kids[4],
                                   uint32_t *nk_out) {
  uint32_t nk = 0;
  if (lbvh_leafish(a, it.ref)) {
    kids[nk++] = it.ref;
  } else {
    kids[nk++] = a.left[it.ref];
    kids[nk++] = a.right[it.ref];
    while (nk < 4) {
      int best = -1;
      float best_area = -1.f;
      for (uint32_t k = 0; k < nk; ++k)
        if (!lbvh_leafish(a, kids[k]) && lbvh_area(a, kids[k]) > best_area) {
          best = int(k);
          best_area = lbvh_area(a, kids[k]);
        }
      if (best < 0)
        break;
      const int32_t open = kids[best];
      kids[best] = a.left[open];
      kids[nk++] = a.right[open];
    }
  }
  uint32_t n_inner = 0;
  for (uint32_t k = 0; k < nk; ++k)
    n_inner += lbvh_leafish(a, kids[k]) ? 0u : 1u;
  *nk_out = nk;
  return n_inner;
}

// The node of a level item: quantises its children's padded boxes, writes
// the triangles of each leaf child at its part of the node's output range
// and gives its inner children the node slots first_slot, first_slot + 1, ...
// (queued in `next` at slot - next_base, with their output ranges). Returns
// false if the quantised node is not conservative.
PT_INLINE bool lbvh_level_node(const LbvhArgs &a, const LbvhRoot &it, const int32_t kids[4],
                               uint32_t nk, uint32_t first_slot, uint32_t next_base,
                               LbvhRoot *next) {
  const float pad = a.scene_box[8 * size_t(it.scene) + 6];
  const uint32_t base = a.tri_begin[it.scene];
  float clo[4][3], chi[4][3];
  int32_t child[4];
  uint32_t pos = it.tri_first; // the kids' triangles follow each other in kid order
  uint32_t slot = first_slot;
  for (uint32_t k = 0; k < nk; ++k) {
    const int32_t ref = kids[k];
    const float *b = ref < 0 ? a.lbox + 6 * size_t(~ref) : a.ibox + 6 * size_t(ref);
    for (int c = 0; c < 3; ++c) {
      clo[k][c] = b[c] - pad;
      chi[k][c] = b[3 + c] + pad;
    }
    if (lbvh_leafish(a, ref)) {
      child[k] = ~int32_t(pos - base);
      lbvh_emit_leaf(a, ref, pos, it.scene);
    } else {
      child[k] = int32_t(slot - it.root_slot);
      next[slot - next_base] = LbvhRoot{ref, slot, it.scene, it.root_slot, pos};
      ++slot;
    }
    pos += lbvh_count(a, ref);
  }
  GpuNode node;
  const bool ok = quantise_gpu_node(clo, chi, nk, child, &node);
  a.nodes[it.slot] = node;
  return ok;
}

} // namespace gpu
} // namespace pt

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>

namespace pt {
namespace gpu {

// Scratch of one build: the tree arrays of LbvhArgs plus the level queues
// (n_internal + n_scenes items each) and the radix sort's temporary storage.
struct LbvhBuffers {
  LbvhArgs a;
  LbvhRoot *frontier[2];
  void *sort_temp;
  size_t sort_temp_bytes;
};

// Bytes of radix-sort temporary storage for n keys.
hipError_t lbvh_sort_temp_bytes(uint32_t n, uint32_t n_scenes, size_t *bytes);
enum LbvhBuilder : int { kLbvhRadixTree = 1, kLbvhPloc = 2 };

// Build the 4-wide trees of every scene with the given builder. roots (host
// array, one per scene with triangles, slots 0..n_roots-1 in order) seed the
// first level.
// *levels receives the depth of the deepest tree + 1; returns
// hipErrorInvalidValue if a tree needs more than kLbvhMaxLevels levels or a
// node fails the quantisation check.
hipError_t build_lbvh(const LbvhBuffers &b, const LbvhRoot *roots, uint32_t n_roots,
                      LbvhBuilder builder, hipStream_t s, uint32_t *levels);

} // namespace gpu
} // namespace pt
#endif
