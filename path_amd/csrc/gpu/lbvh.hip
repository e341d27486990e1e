// GPU LBVH build kernels (see lbvh.h for the algorithm and numbering).
#include "lbvh.h"
#include "radix_sort.h"

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

struct DeviceOps {
  __device__ void fence() { __threadfence(); }
  __device__ uint32_t add(uint32_t *p, uint32_t v) { return atomicAdd(p, v); }
  __device__ void bit_or(uint32_t *p, uint32_t v) { atomicOr(p, v); }
};

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

__global__ __launch_bounds__(256) void k_lbvh_level(LbvhArgs a, const LbvhRoot *items,
                                                    uint32_t n_items, LbvhRoot *next) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n_items)
    return;
  DeviceOps ops;
  lbvh_level_item(a, items[i], next, ops);
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
                      hipStream_t s, uint32_t *levels) {
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
  if (a.n_internal)
    hipLaunchKernelGGL(k_lbvh_karras, dim3((a.n_internal + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_lbvh_leaves, dim3(a.n_scenes), dim3(kLeafBlock), 0, s, a);
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
  int cur = 0;
  while (n > 0) {
    if (level == kLbvhMaxLevels)
      return hipErrorInvalidValue; // deeper than the traversal stack allows
    if ((e = hipMemsetAsync(a.counters + 1, 0, sizeof(uint32_t), s)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_lbvh_level, dim3((n + 255) / 256), dim3(256), 0, s, a,
                       b.frontier[cur], n, b.frontier[cur ^ 1]);
    if ((e = hipGetLastError()) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(host, a.counters, sizeof(host), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess)
      return e;
    n = host[1];
    cur ^= 1;
    ++level;
  }
  *levels = level;
  return host[2] ? hipErrorInvalidValue : hipSuccess;
}

} // namespace gpu
} // namespace pt
