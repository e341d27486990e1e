// GPU LBVH build kernels (see lbvh.h for the algorithm and numbering).
#include "lbvh.h"
#include "radix_sort.h"
#include "wave.h"

#include <vector>

namespace pt {
namespace gpu {

// Per scene: bounds of the triangle-box centres and the padding of the host
// packer (2^-13 of the largest vertex coordinate magnitude, at least 1e-3).
__global__ __launch_bounds__(256) void k_lbvh_scene_box(LbvhArgs a) {
  __shared__ float red[7][256];
  const uint32_t s = blockIdx.x, t0 = a.tri_begin[s], t1 = a.tri_begin[s + 1];
  float v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 1e-3f};
  for (uint32_t t = t0 + threadIdx.x; t < t1; t += 256u) {
    float lo[3], hi[3];
    lbvh_tri_box(a.shade[t], lo, hi);
    for (int k = 0; k < 3; ++k) {
      const float c = (lo[k] + hi[k]) * 0.5f;
      v[k] = fminf(v[k], c);
      v[3 + k] = fmaxf(v[3 + k], c);
      v[6] = fmaxf(v[6], fmaxf(fabsf(lo[k]), fabsf(hi[k])));
    }
  }
  for (int k = 0; k < 7; ++k)
    red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (uint32_t off = 128; off >= 1; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < 7; ++k) {
        const float o = red[k][threadIdx.x + off];
        float &m = red[k][threadIdx.x];
        m = (k >= 3) ? fmaxf(m, o) : fminf(m, o);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6)
    a.scene_box[8 * size_t(s) + threadIdx.x] = red[threadIdx.x][0];
  if (threadIdx.x == 6)
    a.scene_box[8 * size_t(s) + 6] = red[6][0] * 0x1p-13f;
}

__global__ __launch_bounds__(256) void k_lbvh_keys(LbvhArgs a) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= a.n_tris)
    return;
  a.keys[t] = lbvh_key(a, t);
  a.vals[t] = t;
}

__global__ __launch_bounds__(256) void k_lbvh_clear(LbvhArgs a) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < a.n_internal) {
    a.parent_int[i] = -1;
    a.visits[i] = 0;
  }
  if (i < a.n_tris)
    a.parent_leaf[i] = -1;
  if (i < 3)
    a.counters[i] = 0;
}

__global__ __launch_bounds__(256) void k_lbvh_karras(LbvhArgs a) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g < a.n_internal)
    lbvh_karras(a, g);
}

// Synchronisation inside one workgroup only. A device-scope release on this
// GPU writes back the L2 of the XCD (the XCDs' L2s are not coherent with each
// other), which made the bottom-up pass cost ~22 ms for 3 M triangles; when
// every thread that touches a tree runs in the same workgroup, workgroup
// scope is enough and costs a wait for outstanding stores.
struct WorkgroupOps {
  __device__ void fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }
  __device__ uint32_t add(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ void bit_or(uint32_t *p, uint32_t v) {
    (void)__hip_atomic_fetch_or(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
};

// One workgroup per scene: the scene's trees never share nodes with another
// scene (parent_int is -1 at each scene root), so all the atomics and boxes of
// one tree stay inside its workgroup. Threads stride over the scene's sorted
// leaves; a walk that meets a sibling finished by an earlier stride sees its
// box through the acquire-release visit counter.
constexpr uint32_t kLeafBlock = 1024;

__global__ __launch_bounds__(kLeafBlock) void k_lbvh_leaves(LbvhArgs a) {
  const uint32_t begin = a.tri_begin[blockIdx.x], end = a.tri_begin[blockIdx.x + 1];
  WorkgroupOps ops;
  for (uint32_t p = begin + threadIdx.x; p < end; p += kLeafBlock)
    lbvh_leaf(a, p, ops);
}

__global__ __launch_bounds__(256) void k_lbvh_leaf_boxes(LbvhArgs a) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p < a.n_tris)
    lbvh_leaf_box(a, p);
}

// ---------------------------------------------------------------------------
// PLOC, one workgroup per scene. Cluster c of a pass is two float4 records
// (box lo + ref, box hi + triangle count) in Morton order. A pass finds each
// cluster's nearest neighbour within kPlocRadius positions by the surface
// area of the union box, merges mutual pairs into new internal nodes (at the
// lower position; the higher one drops out) and compacts the survivors in
// order into the other buffer. Candidate pairs are ordered by (area, lower
// position, higher position): the least pair of a pass is mutual, so every
// pass merges at least one pair and n - 1 passes are the most a scene of n
// triangles can take (a pass cap guards against NaN boxes, where the areas
// compare false: the build then reports failure and the batch falls back to
// host trees). Internal nodes are numbered as they are created: the merge
// that takes the cluster count from c to c - 1 makes local node c - 2, so
// the last merge makes the root, node 0.
// ---------------------------------------------------------------------------
constexpr uint32_t kPlocBlock = 1024;
constexpr uint32_t kPlocWaves = kPlocBlock / 64;

__device__ __forceinline__ float ploc_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return dx * dy + dy * dz + dz * dx;
}

__global__ __launch_bounds__(kPlocBlock) void k_ploc(LbvhArgs a) {
  __shared__ uint32_t s_cnt[2][kPlocWaves];
  const uint32_t scene = blockIdx.x;
  const uint32_t tb = a.tri_begin[scene], n0 = a.tri_begin[scene + 1] - tb;
  if (n0 < 2)
    return; // no internal nodes (the whole block returns)
  const uint32_t ib = a.int_begin[scene];
  float4 *cur = reinterpret_cast<float4 *>(a.clusters[0]) + 2 * size_t(tb);
  float4 *nxt = reinterpret_cast<float4 *>(a.clusters[1]) + 2 * size_t(tb);
  int32_t *nn = a.nn + tb;
  for (uint32_t p = threadIdx.x; p < n0; p += kPlocBlock) {
    const float *lb = a.lbox + 6 * size_t(tb + p);
    cur[2 * p] = make_float4(lb[0], lb[1], lb[2], __int_as_float(~int32_t(tb + p)));
    cur[2 * p + 1] = make_float4(lb[3], lb[4], lb[5], __uint_as_float(1u));
  }
  __syncthreads();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t n = n0, passes = 0;
  while (n > 1) {
    if (++passes >= n0) { // n0 - 1 passes always suffice (see above)
      if (threadIdx.x == 0)
        atomicOr(&a.counters[2], 2u);
      return;
    }
    // nearest neighbours
    for (uint32_t i = threadIdx.x; i < n; i += kPlocBlock) {
      const float4 lo = cur[2 * i], hi = cur[2 * i + 1];
      float best = INFINITY;
      int32_t bj = -1;
      const uint32_t j0 = i > uint32_t(kPlocRadius) ? i - uint32_t(kPlocRadius) : 0u;
      const uint32_t j1 = min(n - 1, i + uint32_t(kPlocRadius));
      for (uint32_t j = j0; j <= j1; ++j) {
        if (j == i)
          continue;
        const float ar = ploc_area(lo, hi, cur[2 * j], cur[2 * j + 1]);
        // pairs with equal areas: (min, max) of the positions decides; for
        // this i, a candidate below it orders before one above it, and
        // among the lower ones the smaller, among the upper the smaller
        bool better = ar < best;
        if (!better && ar == best && bj >= 0) {
          const uint32_t b = uint32_t(bj);
          const uint32_t jmin = min(i, j), jmax = max(i, j);
          const uint32_t bmin = min(i, b), bmax = max(i, b);
          better = jmin < bmin || (jmin == bmin && jmax < bmax);
        }
        if (better) {
          best = ar;
          bj = int32_t(j);
        }
      }
      nn[i] = bj;
    }
    __syncthreads();
    // merge mutual pairs and compact, 1024 positions at a time
    uint32_t n_keep = 0, n_merge = 0;
    for (uint32_t base = 0; base < n; base += kPlocBlock) {
      const uint32_t i = base + threadIdx.x;
      const int32_t j = i < n ? nn[i] : -1;
      const bool mutual = j >= 0 && nn[j] == int32_t(i);
      const bool merge = mutual && int32_t(i) < j;
      const bool keep = i < n && !(mutual && int32_t(i) > j);
      uint32_t wk, wm;
      const uint32_t pk = ballot_prefix(keep, &wk), pm = ballot_prefix(merge, &wm);
      if (lane_id() == 0) {
        s_cnt[0][wave] = wk;
        s_cnt[1][wave] = wm;
      }
      __syncthreads();
      uint32_t bk = 0, bm = 0, tk = 0, tm = 0;
      for (uint32_t w = 0; w < kPlocWaves; ++w) {
        const uint32_t x = s_cnt[0][w], y = s_cnt[1][w];
        bk += w < wave ? x : 0u;
        bm += w < wave ? y : 0u;
        tk += x;
        tm += y;
      }
      if (keep) {
        float4 lo = cur[2 * i], hi = cur[2 * i + 1];
        if (merge) {
          const float4 lo2 = cur[2 * j], hi2 = cur[2 * j + 1];
          const uint32_t g = ib + (n - 2u - (n_merge + bm + pm));
          const uint32_t cnt = __float_as_uint(hi.w) + __float_as_uint(hi2.w);
          a.left[g] = __float_as_int(lo.w);
          a.right[g] = __float_as_int(lo2.w);
          a.count[g] = cnt;
          lo = make_float4(fminf(lo.x, lo2.x), fminf(lo.y, lo2.y), fminf(lo.z, lo2.z),
                           __int_as_float(int32_t(g)));
          hi = make_float4(fmaxf(hi.x, hi2.x), fmaxf(hi.y, hi2.y), fmaxf(hi.z, hi2.z),
                           __uint_as_float(cnt));
          float *ob = a.ibox + 6 * size_t(g);
          ob[0] = lo.x, ob[1] = lo.y, ob[2] = lo.z;
          ob[3] = hi.x, ob[4] = hi.y, ob[5] = hi.z;
        }
        const uint32_t o = n_keep + bk + pk;
        nxt[2 * o] = lo;
        nxt[2 * o + 1] = hi;
      }
      n_keep += tk;
      n_merge += tm;
      __syncthreads(); // s_cnt is rewritten by the next chunk; nxt complete
    }
    n = n_keep;
    float4 *t = cur;
    cur = nxt;
    nxt = t;
  }
}

// One thread per item of a level. The slots of the next level's nodes are
// handed out with one device atomic per wave (a wave prefix sum of the
// threads' inner-child counts): one atomic per inner child on the same
// counter cost ~119 us per level on 1024 scenes. A level's slots are
// contiguous from next_base, so a node's slot also gives its queue position.
__global__ __launch_bounds__(256) void k_lbvh_level(LbvhArgs a, const LbvhRoot *items,
                                                    uint32_t n_items, uint32_t next_base,
                                                    LbvhRoot *next) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const bool valid = i < n_items;
  LbvhRoot it{};
  int32_t kids[4] = {0, 0, 0, 0};
  uint32_t nk = 0, n_inner = 0;
  if (valid) {
    it = items[i];
    n_inner = lbvh_level_kids(a, it, kids, &nk);
  }
  const uint32_t lane = lane_id();
  uint32_t incl = n_inner;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t v = uint32_t(__shfl_up(int(incl), off));
    if (lane >= off)
      incl += v;
  }
  const uint32_t total = uint32_t(__shfl(int(incl), 63));
  uint32_t wave_base = 0;
  if (lane == 0 && total)
    wave_base = atomicAdd(&a.counters[0], total);
  wave_base = uint32_t(__shfl(int(wave_base), 0));
  if (valid && !lbvh_level_node(a, it, kids, nk, wave_base + incl - n_inner, next_base, next))
    atomicOr(&a.counters[2], 1u);
}

static uint32_t key_end_bit(uint32_t n_scenes) {
  uint32_t bits = 1;
  while (bits < 32 && (1u << bits) < n_scenes)
    ++bits;
  return 32 + bits;
}

hipError_t lbvh_sort_temp_bytes(uint32_t n, uint32_t n_scenes, size_t *bytes) {
  (void)n_scenes;
  *bytes = radix_sort_temp_bytes(n);
  return hipSuccess;
}

hipError_t build_lbvh(const LbvhBuffers &b, const LbvhRoot *roots, uint32_t n_roots,
                      LbvhBuilder builder, hipStream_t s, uint32_t *levels) {
  const LbvhArgs &a = b.a;
  *levels = 0;
  if (a.n_tris == 0 || n_roots == 0)
    return hipSuccess;
  hipError_t e;
  const uint32_t span = a.n_tris > a.n_internal ? a.n_tris : a.n_internal;
  hipLaunchKernelGGL(k_lbvh_clear, dim3((span + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_lbvh_scene_box, dim3(a.n_scenes), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_lbvh_keys, dim3((a.n_tris + 255) / 256), dim3(256), 0, s, a);
  if ((e = hipGetLastError()) != hipSuccess)
    return e;
  // (scene, Morton code) keys; equal keys keep triangle order
  if ((e = radix_sort_pairs(a.keys, a.vals, a.keys_sorted, a.vals_sorted, a.n_tris,
                            key_end_bit(a.n_scenes), b.sort_temp, b.sort_temp_bytes, s)) !=
      hipSuccess)
    return e;
  if (builder == kLbvhPloc) {
    hipLaunchKernelGGL(k_lbvh_leaf_boxes, dim3((a.n_tris + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_ploc, dim3(a.n_scenes), dim3(kPlocBlock), 0, s, a);
  } else {
    if (a.n_internal)
      hipLaunchKernelGGL(k_lbvh_karras, dim3((a.n_internal + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_lbvh_leaves, dim3(a.n_scenes), dim3(kLeafBlock), 0, s, a);
  }
  if ((e = hipGetLastError()) != hipSuccess)
    return e;
  // levels: roots take slots 0..n_roots-1, the others are allocated in order
  if ((e = hipMemcpyAsync(b.frontier[0], roots, sizeof(LbvhRoot) * n_roots,
                          hipMemcpyHostToDevice, s)) != hipSuccess)
    return e;
  const uint32_t first_free[1] = {n_roots};
  if ((e = hipMemcpyAsync(a.counters, first_free, sizeof(uint32_t), hipMemcpyHostToDevice,
                          s)) != hipSuccess)
    return e;
  uint32_t n = n_roots, level = 0, host[3] = {0, 0, 0};
  uint32_t next_base = n_roots; // first slot of the next level's nodes
  int cur = 0;
  while (n > 0) {
    if (level == kLbvhMaxLevels)
      return hipErrorInvalidValue; // deeper than the traversal stack allows
    hipLaunchKernelGGL(k_lbvh_level, dim3((n + 255) / 256), dim3(256), 0, s, a,
                       b.frontier[cur], n, next_base, b.frontier[cur ^ 1]);
    if ((e = hipGetLastError()) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(host, a.counters, sizeof(host), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess)
      return e;
    n = host[0] - next_base; // slots handed out by this level
    next_base = host[0];
    cur ^= 1;
    ++level;
  }
  *levels = level;
  return host[2] ? hipErrorInvalidValue : hipSuccess;
}

} // namespace gpu
} // namespace pt
