// Batched multi-scene rendering (training-data generation) and the launch
// autotuner.
#include "renderer.h"
#include "scene_gen.h"
#include "settings.h"

#include "../gpu/trace.h"
#include "../gpu/wavefront.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>
#include <type_traits>
#include <vector>

using namespace pt;
using namespace pt::host;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error(PT_ERR_HIP, std::string(#expr) + ": " +                 \
                                       hipGetErrorString(e_));                 \
  } while (0)

namespace {

// Run f(i) for i in [0, n) on up to 16 host threads.
template <class F> void parallel_for(size_t n, F f) {
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  unsigned nt = unsigned(std::min<size_t>(hw, n));
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i)
      f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      for (size_t k = next++; k < n; k = next++)
        f(k);
    });
  for (auto &th : pool)
    th.join();
}

// Flatten, build and pack the BVHs of the distinct scenes in parallel (host
// trees; with GPU trees pack() writes the flattened arrays itself).
void build_all(const pt_batch_item *items, uint32_t n) {
  std::vector<const pt_scene *> uniq;
  for (uint32_t i = 0; i < n; ++i)
    uniq.push_back(items[i].scene);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  parallel_for(uniq.size(), [&](size_t k) { uniq[k]->packed(); });
}

// Byte layout of the concatenated scene arrays in the pinned staging buffer
// (each section 256-byte aligned).
struct Layout {
  enum { kNodes, kTris, kShade, kMats, kLightTris, kLightCdf, kDesc, kBegin, kN };
  size_t off[kN + 1];
  size_t count[kN];
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// gpu_bvh: the trees are built on the device after the upload; scene i's
// root takes node slot (its rank among the scenes with triangles) and its
// leaf triangles sit at its shading offset. The shading records, materials
// and emitters are then written straight into the pinned staging buffer from
// the scene (the values flatten() gives), without the scene's cached
// flattened arrays: no per-scene allocation and no second copy.
int pack(pt_renderer &r, const pt_batch_item *items, uint32_t n, uint32_t spp, bool gpu_bvh,
         std::vector<gpu::SceneDesc> *desc, Layout *lay, uint64_t *n_pixels,
         uint64_t *total_samples) {
  desc->resize(n);
  std::vector<unsigned long long> begin(n + 1, 0);
  size_t c[Layout::kN] = {};
  uint64_t px_total = 0;
  uint32_t roots = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const pt_batch_item &it = items[i];
    const pt_scene &sc = *it.scene;
    size_t n_tris, n_lights;
    const PackedScene *pk = nullptr;
    if (gpu_bvh) {
      n_tris = flat_triangle_count(sc.scene);
      n_lights = flat_light_count(sc.scene);
    } else {
      const FlatScene &f = sc.flat();
      pk = &sc.packed();
      if (pk->bvh.stack_need > uint32_t(gpu_stack_size()))
        return set_error(PT_ERR_UNSUPPORTED, "batch item " + std::to_string(i) +
                                                 ": BVH too deep for the traversal stack");
      n_tris = f.tris.size();
      n_lights = f.light_tris.size();
    }
    gpu::SceneDesc d{};
    d.node_off = uint32_t(c[Layout::kNodes]);
    d.tri_off = uint32_t(c[Layout::kTris]);
    d.shade_off = uint32_t(c[Layout::kShade]);
    d.mat_off = uint32_t(c[Layout::kMats]);
    d.light_off = uint32_t(c[Layout::kLightTris]);
    d.n_lights = uint32_t(n_lights);
    d.x_dim = it.x_dim;
    d.y_dim = it.y_dim;
    d.pixel_off = uint32_t(px_total);
    d.has_nodes = n_tris ? 1u : 0u;
    if (gpu_bvh) {
      d.node_off = roots;
      d.tri_off = d.shade_off;
      roots += d.has_nodes;
    }
    std::memcpy(d.film_to_world, it.film_to_world, sizeof(d.film_to_world));
    (*desc)[i] = d;
    if (!gpu_bvh) {
      c[Layout::kNodes] += pk->bvh.nodes.size();
      c[Layout::kTris] += pk->bvh.tris.size();
    }
    c[Layout::kShade] += n_tris;
    c[Layout::kMats] += sc.scene.materials.size();
    c[Layout::kLightTris] += n_lights;
    c[Layout::kLightCdf] += n_lights;
    uint64_t px = uint64_t(it.x_dim) * it.y_dim;
    px_total += px;
    begin[i + 1] = begin[i] + px * spp;
    if (px_total >= 0xffffffffull || c[Layout::kNodes] >= 0x7fffffffull ||
        c[Layout::kTris] >= 0x7fffffffull)
      return set_error(PT_ERR_INVALID, "batch too large for 32-bit offsets");
  }
  c[Layout::kDesc] = n;
  c[Layout::kBegin] = n + 1;
  const size_t elem[Layout::kN] = {sizeof(GpuNode),         sizeof(GpuTri),
                                   sizeof(ShadeTri<float>), sizeof(Material<float>),
                                   sizeof(TriGeom<float>),  sizeof(float),
                                   sizeof(gpu::SceneDesc),  sizeof(unsigned long long)};
  lay->off[0] = 0;
  for (int k = 0; k < Layout::kN; ++k) {
    lay->count[k] = c[k];
    lay->off[k + 1] = lay->off[k] + align256(c[k] * elem[k]);
  }
  hipError_t e = r.b_stage_in.ensure(lay->off[Layout::kN]);
  if (e != hipSuccess)
    return set_error(PT_ERR_HIP, std::string("pinned staging: ") + hipGetErrorString(e));
  char *base = static_cast<char *>(r.b_stage_in.ptr);
  std::memcpy(base + lay->off[Layout::kDesc], desc->data(), n * sizeof(gpu::SceneDesc));
  std::memcpy(base + lay->off[Layout::kBegin], begin.data(), (n + 1) * sizeof(begin[0]));
  auto at = [&](int sec, size_t first) { return base + lay->off[sec] + first * elem[sec]; };
  if (gpu_bvh) {
    // Scenes are written in order by 16 threads; the calling thread also
    // sends each finished prefix of scenes to the device (at least 16 MB of
    // shading records at a time, or the rest), so the PCIe upload runs while
    // later scenes are still being written.
    hipError_t e = hipSuccess;
    auto grow = [&](auto &buf, int sec) {
      if (e == hipSuccess)
        e = buf.ensure(std::max<size_t>(1, c[sec]));
    };
    grow(r.b_shade, Layout::kShade);
    grow(r.b_mats, Layout::kMats);
    grow(r.b_light_tris, Layout::kLightTris);
    grow(r.b_light_cdf, Layout::kLightCdf);
    if (e != hipSuccess)
      return set_error(PT_ERR_HIP, std::string("batch arrays: ") + hipGetErrorString(e));
    auto fill = [&](size_t i) {
      const gpu::SceneDesc &d = (*desc)[i];
      const Scene &scene = items[i].scene->scene;
      write_shade_tris(scene, reinterpret_cast<ShadeTri<float> *>(at(Layout::kShade, d.shade_off)));
      auto *m = reinterpret_cast<Material<float> *>(at(Layout::kMats, d.mat_off));
      for (size_t k = 0; k < scene.materials.size(); ++k)
        m[k] = scene.materials[k].to<float>();
      flatten_lights(scene, reinterpret_cast<TriGeom<float> *>(at(Layout::kLightTris, d.light_off)),
                     reinterpret_cast<float *>(at(Layout::kLightCdf, d.light_off)));
    };
    std::unique_ptr<std::atomic<uint8_t>[]> done(new std::atomic<uint8_t>[n]());
    std::atomic<size_t> next{0};
    size_t sent = 0; // scenes whose arrays have been queued for upload
    auto send = [&](size_t upto) {
      const gpu::SceneDesc &a = (*desc)[sent];
      auto copy = [&](auto &buf, int sec, size_t from, size_t to) {
        if (e == hipSuccess && to > from)
          e = hipMemcpyAsync(buf.ptr + from, at(sec, from), (to - from) * elem[sec],
                             hipMemcpyHostToDevice, r.stream);
      };
      const bool last = upto == n;
      const gpu::SceneDesc *b = last ? nullptr : &(*desc)[upto];
      copy(r.b_shade, Layout::kShade, a.shade_off, last ? c[Layout::kShade] : b->shade_off);
      copy(r.b_mats, Layout::kMats, a.mat_off, last ? c[Layout::kMats] : b->mat_off);
      copy(r.b_light_tris, Layout::kLightTris, a.light_off,
           last ? c[Layout::kLightTris] : b->light_off);
      copy(r.b_light_cdf, Layout::kLightCdf, a.light_off,
           last ? c[Layout::kLightCdf] : b->light_off);
      sent = upto;
    };
    const size_t min_send = (size_t(16) << 20) / sizeof(ShadeTri<float>);
    auto worker = [&](bool sender) {
      for (size_t i = next++; i < n; i = next++) {
        fill(i);
        done[i].store(1, std::memory_order_release);
        if (!sender)
          continue;
        size_t p = sent;
        while (p < n && done[p].load(std::memory_order_acquire))
          ++p;
        if (p > sent && (p == n || (*desc)[p].shade_off - (*desc)[sent].shade_off >= min_send))
          send(p);
      }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < std::min<size_t>(hw, n); ++t)
      pool.emplace_back(worker, false);
    worker(true);
    for (auto &th : pool)
      th.join();
    if (sent < n)
      send(n);
    if (e != hipSuccess)
      return set_error(PT_ERR_HIP, std::string("batch upload: ") + hipGetErrorString(e));
  } else {
    parallel_for(n, [&](size_t i) {
      const gpu::SceneDesc &d = (*desc)[i];
      const FlatScene &f = items[i].scene->flat();
      const PackedScene &pk = items[i].scene->packed();
      auto put = [&](int sec, size_t first, const void *src, size_t bytes) {
        if (bytes)
          std::memcpy(at(sec, first), src, bytes);
      };
      put(Layout::kNodes, d.node_off, pk.bvh.nodes.data(), pk.bvh.nodes.size() * sizeof(GpuNode));
      put(Layout::kTris, d.tri_off, pk.bvh.tris.data(), pk.bvh.tris.size() * sizeof(GpuTri));
      put(Layout::kShade, d.shade_off, pk.shade.data(), pk.shade.size() * sizeof(ShadeTri<float>));
      put(Layout::kMats, d.mat_off, pk.mats.data(), pk.mats.size() * sizeof(Material<float>));
      put(Layout::kLightTris, d.light_off, f.light_tris.data(),
          f.light_tris.size() * sizeof(TriGeom<float>));
      put(Layout::kLightCdf, d.light_off, f.light_cdf.data(), f.light_cdf.size() * sizeof(float));
    });
  }
  *n_pixels = px_total;
  *total_samples = begin[n];
  return PT_OK;
}

// Trees of every scene of the batch on the GPU (gpu/lbvh.h, r.gpu_builder). Returns
// PT_OK with *fallback set if some tree is too deep for the traversal stack
// (or fails the quantisation check): the caller then builds on the host.
int build_gpu_trees(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                    const std::vector<gpu::SceneDesc> &desc, hipStream_t s, bool *fallback,
                    uint32_t *levels) {
  *fallback = false;
  std::vector<uint32_t> tri_begin(n + 1, 0), int_begin(n + 1, 0);
  std::vector<gpu::LbvhRoot> roots;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t nt = uint32_t(flat_triangle_count(items[i].scene->scene));
    tri_begin[i + 1] = tri_begin[i] + nt;
    int_begin[i + 1] = int_begin[i] + (nt >= 2 ? nt - 1 : 0);
    if (desc[i].has_nodes != (nt > 0 ? 1u : 0u) || (nt && desc[i].tri_off != tri_begin[i]))
      return set_error(PT_ERR_INVALID, "batch layout does not match the GPU tree build");
    if (nt)
      roots.push_back({nt >= 2 ? int32_t(int_begin[i]) : ~int32_t(tri_begin[i]),
                       desc[i].node_off, i, desc[i].node_off, tri_begin[i]});
  }
  const uint32_t nt = tri_begin[n], ni = int_begin[n];
  *levels = 0;
  if (nt == 0)
    return PT_OK;
  const size_t one = 1;
  HIP_TRY(r.lb_tri_begin.upload(tri_begin.data(), n + 1, s));
  HIP_TRY(r.lb_int_begin.upload(int_begin.data(), n + 1, s));
  HIP_TRY(r.lb_scene_box.ensure(size_t(n) * 8));
  HIP_TRY(r.lb_keys.ensure(nt));
  HIP_TRY(r.lb_keys_sorted.ensure(nt));
  HIP_TRY(r.lb_vals.ensure(nt));
  HIP_TRY(r.lb_vals_sorted.ensure(nt));
  HIP_TRY(r.lb_left.ensure(std::max(one, size_t(ni))));
  HIP_TRY(r.lb_right.ensure(std::max(one, size_t(ni))));
  HIP_TRY(r.lb_count.ensure(std::max(one, size_t(ni))));
  uint32_t max_tris = 0;
  for (uint32_t i = 0; i < n; ++i)
    max_tris = std::max(max_tris, tri_begin[i + 1] - tri_begin[i]);
  const bool ploc = r.gpu_builder == gpu::kLbvhPloc && max_tris <= gpu::kPlocMaxSceneTris;
  if (ploc) {
    HIP_TRY(r.lb_clusters[0].ensure(8 * size_t(nt)));
    HIP_TRY(r.lb_clusters[1].ensure(8 * size_t(nt)));
    HIP_TRY(r.lb_nn.ensure(nt));
  }
  HIP_TRY(r.lb_parent_int.ensure(std::max(one, size_t(ni))));
  HIP_TRY(r.lb_visits.ensure(std::max(one, size_t(ni))));
  HIP_TRY(r.lb_ibox.ensure(std::max(one, size_t(ni)) * 6));
  HIP_TRY(r.lb_parent_leaf.ensure(nt));
  HIP_TRY(r.lb_lbox.ensure(size_t(nt) * 6));
  HIP_TRY(r.lb_counters.ensure(4));
  // every 4-wide node is a root or an inner binary node
  const size_t max_nodes = size_t(ni) + n;
  HIP_TRY(r.lb_frontier[0].ensure(max_nodes));
  HIP_TRY(r.lb_frontier[1].ensure(max_nodes));
  HIP_TRY(r.b_nodes.ensure(max_nodes));
  HIP_TRY(r.b_tris.ensure(nt));
  size_t temp = 0;
  HIP_TRY(gpu::lbvh_sort_temp_bytes(nt, n, &temp));
  HIP_TRY(r.lb_sort_temp.ensure(std::max(one, temp)));
  gpu::LbvhBuffers b{};
  gpu::LbvhArgs &a = b.a;
  a.shade = r.b_shade.ptr;
  a.tri_begin = r.lb_tri_begin.ptr;
  a.int_begin = r.lb_int_begin.ptr;
  a.n_scenes = n;
  a.n_tris = nt;
  a.n_internal = ni;
  a.scene_box = r.lb_scene_box.ptr;
  a.keys = r.lb_keys.ptr;
  a.keys_sorted = r.lb_keys_sorted.ptr;
  a.vals = r.lb_vals.ptr;
  a.vals_sorted = r.lb_vals_sorted.ptr;
  a.left = r.lb_left.ptr;
  a.right = r.lb_right.ptr;
  a.count = r.lb_count.ptr;
  a.clusters[0] = r.lb_clusters[0].ptr;
  a.clusters[1] = r.lb_clusters[1].ptr;
  a.nn = r.lb_nn.ptr;
  a.parent_int = r.lb_parent_int.ptr;
  a.parent_leaf = r.lb_parent_leaf.ptr;
  a.ibox = r.lb_ibox.ptr;
  a.lbox = r.lb_lbox.ptr;
  a.visits = r.lb_visits.ptr;
  a.nodes = r.b_nodes.ptr;
  a.tris = r.b_tris.ptr;
  a.counters = r.lb_counters.ptr;
  b.frontier[0] = r.lb_frontier[0].ptr;
  b.frontier[1] = r.lb_frontier[1].ptr;
  b.sort_temp = r.lb_sort_temp.ptr;
  b.sort_temp_bytes = temp;
  const hipError_t e =
      gpu::build_lbvh(b, roots.data(), uint32_t(roots.size()),
                      ploc ? gpu::kLbvhPloc : gpu::kLbvhRadixTree, s, levels);
  if (e == hipErrorInvalidValue) {
    *fallback = true;
    return PT_OK;
  }
  HIP_TRY(e);
  return PT_OK;
}

// Make the batch's scenes resident in the renderer's concatenated arrays
// (b_nodes, b_tris, b_shade, b_mats, b_light_*, b_desc, b_sample_begin).
// The scenes of the last batch stay on the GPU: the same scenes (by
// generation), sizes, spp and tree builder skip the host packing, the upload
// and the tree builds; new cameras only re-upload the descriptors. When a GPU
// tree build falls back, r.gpu_bvh is cleared and nothing is uploaded: the
// caller repeats the call with host trees (and restores r.gpu_bvh).
int ensure_batch_scenes(pt_renderer &r, const pt_batch_item *items, uint32_t n, uint32_t spp,
                        bool gpu_bvh, const pt_settings &settings,
                        std::vector<gpu::SceneDesc> *desc_out, uint64_t *n_pixels_out,
                        uint64_t *total_out) {
  if (!gpu_bvh)
    for (uint32_t i = 0; i < n; ++i)
      items[i].scene->use_tree_settings(tree_settings(settings));
  std::vector<uint64_t> key;
  key.reserve(3 * size_t(n) + 2);
  key.push_back(spp);
  key.push_back(gpu_bvh ? uint64_t(r.gpu_builder) : 0u);
  for (uint32_t i = 0; i < n; ++i) {
    key.push_back(items[i].scene->generation);
    key.push_back(items[i].scene->mat_generation);
    key.push_back(uint64_t(items[i].x_dim) | (uint64_t(items[i].y_dim) << 32));
  }
  const bool resident = key == r.batch_key;
  std::vector<gpu::SceneDesc> &desc = *desc_out;
  Layout lay;
  uint64_t &n_pixels = *n_pixels_out, &total = *total_out;
  int rc = PT_OK;
  hipStream_t s = r.stream;
  if (resident) {
    desc = r.batch_desc;
    n_pixels = r.batch_pixels;
    total = r.batch_samples;
    // The cameras are not part of the key: a resident batch rendered from
    // new viewpoints re-uploads only the per-scene descriptors (96 B each).
    bool moved = false;
    for (uint32_t i = 0; i < n; ++i)
      if (std::memcmp(desc[i].film_to_world, items[i].film_to_world,
                      sizeof(desc[i].film_to_world)) != 0) {
        std::memcpy(desc[i].film_to_world, items[i].film_to_world,
                    sizeof(desc[i].film_to_world));
        moved = true;
      }
    if (moved) {
      HIP_TRY(hipMemcpyAsync(r.b_desc.ptr, desc.data(), n * sizeof(gpu::SceneDesc),
                             hipMemcpyHostToDevice, s));
      HIP_TRY(hipStreamSynchronize(s));
      r.batch_desc = desc;
    }
  } else {
  r.batch_key.clear();
  using clock = std::chrono::steady_clock;
  auto ms_since = [](clock::time_point t) {
    return std::chrono::duration<double, std::milli>(clock::now() - t).count();
  };
  auto t_phase = clock::now();
  if (!gpu_bvh)
    build_all(items, n);
  r.last_batch_phase_ms[0] = ms_since(t_phase);
  t_phase = clock::now();
  rc = pack(r, items, n, spp, gpu_bvh, &desc, &lay, &n_pixels, &total);
  if (rc != PT_OK)
    return rc;
  r.last_batch_phase_ms[1] = ms_since(t_phase);
  t_phase = clock::now();
  {
    const char *base = static_cast<const char *>(r.b_stage_in.ptr);
    auto up = [&](auto &buf, int sec) {
      using T = std::remove_pointer_t<decltype(buf.ptr)>;
      return buf.upload(reinterpret_cast<const T *>(base + lay.off[sec]), lay.count[sec], s);
    };
    if (!gpu_bvh) { // with GPU trees pack() has queued these already
      HIP_TRY(up(r.b_nodes, Layout::kNodes));
      HIP_TRY(up(r.b_tris, Layout::kTris));
      HIP_TRY(up(r.b_shade, Layout::kShade));
      HIP_TRY(up(r.b_mats, Layout::kMats));
      HIP_TRY(up(r.b_light_tris, Layout::kLightTris));
      HIP_TRY(up(r.b_light_cdf, Layout::kLightCdf));
    }
    HIP_TRY(up(r.b_desc, Layout::kDesc));
    HIP_TRY(up(r.b_sample_begin, Layout::kBegin));
    HIP_TRY(hipStreamSynchronize(s));
  }
  r.last_batch_phase_ms[2] = ms_since(t_phase);
  r.last_batch_phase_ms[3] = 0.0;
  if (gpu_bvh) {
    bool fallback = false;
    uint32_t levels = 0;
    const auto tb = std::chrono::steady_clock::now();
    rc = build_gpu_trees(r, items, n, desc, s, &fallback, &levels);
    if (rc != PT_OK)
      return rc;
    HIP_TRY(hipStreamSynchronize(s));
    if (fallback) {
      // a tree too deep for the traversal stack: the caller builds this
      // batch on the host (r.gpu_bvh is cleared for that call)
      r.gpu_bvh = false;
      return PT_OK;
    }
    r.last_gpu_bvh_ms = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - tb).count();
    r.last_gpu_bvh_levels = levels;
    r.last_batch_phase_ms[3] = r.last_gpu_bvh_ms;
  }
  HIP_TRY(hipStreamSynchronize(s));
  r.batch_key = std::move(key);
  r.batch_desc = desc;
  r.batch_pixels = n_pixels;
  r.batch_samples = total;
  }

  return PT_OK;
}

} // namespace

namespace pt {
namespace host {

int render_batch(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                 uint32_t spp, const pt_settings &settings, double *seconds) {
  if (n == 0)
    return PT_OK;
  if (spp == 0)
    return set_error(PT_ERR_INVALID, "samples_per must be positive");
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(r.device));
  for (uint32_t i = 0; i < n; ++i)
    if (!items[i].scene || items[i].x_dim == 0 || items[i].y_dim == 0 || !items[i].out_rgb)
      return set_error(PT_ERR_INVALID, "batch item " + std::to_string(i) +
                                           ": null scene/output or empty image");
  // outputs: all host memory, or all device memory of this GPU
  bool device_out = false;
  for (uint32_t i = 0; i < n; ++i) {
    for (const void *ptr : {static_cast<const void *>(items[i].out_rgb),
                            static_cast<const void *>(items[i].out_gbuffer)}) {
      if (!ptr)
        continue;
      hipPointerAttribute_t attr{};
      bool dev = hipPointerGetAttributes(&attr, ptr) == hipSuccess &&
                 attr.type == hipMemoryTypeDevice;
      if (!dev)
        (void)hipGetLastError();
      if (dev && attr.device != r.device)
        return set_error(PT_ERR_INVALID, "batch item " + std::to_string(i) +
                                             ": output on another GPU");
      if (i == 0 && ptr == items[0].out_rgb)
        device_out = dev;
      else if (dev != device_out)
        return set_error(PT_ERR_INVALID, "batch outputs must be all host or all device memory");
    }
  }
  const bool gpu_bvh = r.gpu_bvh;
  hipStream_t s = r.stream;
  std::vector<gpu::SceneDesc> desc;
  uint64_t n_pixels = 0, total = 0;
  int rc = ensure_batch_scenes(r, items, n, spp, gpu_bvh, settings, &desc, &n_pixels, &total);
  if (rc != PT_OK)
    return rc;
  if (gpu_bvh && !r.gpu_bvh) {
    // a tree too deep for the traversal stack: this batch was built on the host
    rc = render_batch(r, items, n, spp, settings, seconds);
    r.gpu_bvh = true;
    return rc;
  }
  gpu::RngParams rng{};
  rng.type = settings.rng;
  rng.samples_per = spp;
  if (settings.rng == PT_RNG_SOBEL) {
    rc = r.ensure_sobol(settings.max_sample_size, spp);
    if (rc != PT_OK)
      return rc;
    rng.vals = r.sobol_vals.ptr;
    rng.dim_bound = settings.max_sample_size;
    rng.initial_dim_bound = settings.max_sample_size < 64 ? settings.max_sample_size : 64;
  }
  uint64_t slots = settings.max_num_samples_per_launch;
  if (slots == 0)
    return set_error(PT_ERR_INVALID, "max_num_samples_per_launch must be positive");
  slots = std::min<uint64_t>(slots, total);
  const uint32_t trace_blocks = uint32_t(r.n_cus) * r.trace_blocks_for(true);
  gpu::WfParams p{};
  p.settings = integrator_settings(settings);
  p.rng = rng;
  p.batch.scenes = r.b_desc.ptr;
  p.batch.sample_begin = r.b_sample_begin.ptr;
  p.batch.n_scenes = n;
  p.batch.nodes = reinterpret_cast<const float4 *>(r.b_nodes.ptr);
  p.batch.tris = reinterpret_cast<const float4 *>(r.b_tris.ptr);
  p.batch.shade = r.b_shade.ptr;
  p.batch.materials = r.b_mats.ptr;
  p.batch.light_tris = r.b_light_tris.ptr;
  p.batch.light_cdf = r.b_light_cdf.ptr;
  bool want_gbuffer = false;
  for (uint32_t i = 0; i < n; ++i)
    want_gbuffer |= items[i].out_gbuffer != nullptr;
  if (want_gbuffer) {
    HIP_TRY(r.b_gbuffer.ensure(size_t(n_pixels) * 12));
    HIP_TRY(hipMemsetAsync(r.b_gbuffer.ptr, 0, size_t(n_pixels) * 12 * sizeof(float), s));
    p.gbuffer = r.b_gbuffer.ptr;
  }
  WfPipes w;
  rc = wavefront_alloc(r, w, p, uint32_t(slots), total, true, trace_blocks);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(r.rgb.ensure(size_t(n_pixels) * 3));
  r.stats = pt_render_stats{};
  r.stats.profiled = r.profiling ? 1u : 0u;
  HIP_TRY(hipEventRecord(r.ev_start, s));
  rc = wavefront_run(r, w, true, trace_blocks, r.iters_per_sync);
  if (rc != PT_OK)
    return rc;
  // per-scene reduction in the single-scene order
  for (uint32_t i = 0; i < n; ++i) {
    const pt_batch_item &it = items[i];
    WorkDivision wd;
    if (!WorkDivision::make(work_division_settings(settings.render_work_division), spp,
                            it.x_dim, it.y_dim, &wd) ||
        wd.block_size > 1024)
      return set_error(PT_ERR_INVALID, "invalid render work division settings");
    size_t off = desc[i].pixel_off;
    size_t px = size_t(it.x_dim) * it.y_dim;
    if (wd.num_sample_blocks == 1) {
      HIP_TRY(gpu::launch_wf_reduce(r.wf_samples.ptr + 3 * off * spp, wd, spp,
                                    r.rgb.ptr + 3 * off, s));
    } else {
      HIP_TRY(r.rgb2.ensure(px * 3 * wd.num_sample_blocks));
      HIP_TRY(gpu::launch_wf_reduce(r.wf_samples.ptr + 3 * off * spp, wd, spp,
                                    r.rgb2.ptr, s));
      HIP_TRY(gpu::launch_reduce_sample_blocks(r.rgb2.ptr, wd.num_sample_blocks,
                                               uint32_t(px), spp, r.rgb.ptr + 3 * off, s));
    }
  }
  HIP_TRY(hipEventRecord(r.ev_stop, s));
  const size_t rgb_bytes = size_t(n_pixels) * 3 * sizeof(float);
  const size_t gb_bytes = want_gbuffer ? size_t(n_pixels) * 12 * sizeof(float) : 0;
  if (device_out) {
    // results stay in HBM: one copy per array when the caller's buffers are
    // laid out in batch order, else one per scene
    bool packed = true;
    for (uint32_t i = 0; i < n && packed; ++i) {
      const size_t off = desc[i].pixel_off;
      packed = items[i].out_rgb == items[0].out_rgb + 3 * off &&
               (!want_gbuffer || items[i].out_gbuffer == items[0].out_gbuffer + 12 * off);
    }
    if (packed) {
      HIP_TRY(hipMemcpyAsync(items[0].out_rgb, r.rgb.ptr, rgb_bytes, hipMemcpyDeviceToDevice, s));
      if (want_gbuffer)
        HIP_TRY(hipMemcpyAsync(items[0].out_gbuffer, r.b_gbuffer.ptr, gb_bytes,
                               hipMemcpyDeviceToDevice, s));
    } else {
      for (uint32_t i = 0; i < n; ++i) {
        const size_t off = desc[i].pixel_off;
        const size_t px = size_t(items[i].x_dim) * items[i].y_dim;
        HIP_TRY(hipMemcpyAsync(items[i].out_rgb, r.rgb.ptr + 3 * off, px * 12,
                               hipMemcpyDeviceToDevice, s));
        if (items[i].out_gbuffer)
          HIP_TRY(hipMemcpyAsync(items[i].out_gbuffer, r.b_gbuffer.ptr + 12 * off, px * 48,
                                 hipMemcpyDeviceToDevice, s));
      }
    }
    HIP_TRY(hipStreamSynchronize(s));
  } else {
  // results: one copy each into pinned staging, then a parallel scatter
  HIP_TRY(r.b_stage_out.ensure(rgb_bytes + gb_bytes));
  char *stage = static_cast<char *>(r.b_stage_out.ptr);
  HIP_TRY(hipMemcpyAsync(stage, r.rgb.ptr, rgb_bytes, hipMemcpyDeviceToHost, s));
  if (want_gbuffer)
    HIP_TRY(hipMemcpyAsync(stage + rgb_bytes, r.b_gbuffer.ptr, gb_bytes,
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  parallel_for(n, [&](size_t i) {
    size_t off = desc[i].pixel_off;
    size_t px = size_t(items[i].x_dim) * items[i].y_dim;
    std::memcpy(items[i].out_rgb, stage + off * 12, px * 12);
    if (items[i].out_gbuffer)
      std::memcpy(items[i].out_gbuffer, stage + rgb_bytes + off * 48, px * 48);
  });
  }
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
  r.stats.kernel_ms = ms;
  r.stats.samples = total;
  r.stats.rays = r.stats.ext_rays + r.stats.shadow_rays;
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}

int render_rays_batch(pt_renderer &r, const pt_ray_batch_item *items, uint32_t n,
                      uint32_t spp, uint32_t n_steps, const pt_settings &settings,
                      double *seconds) {
  if (spp == 0)
    return set_error(PT_ERR_INVALID, "samples_per must be positive");
  if (n_steps > gpu::kMaxOutputSteps)
    return set_error(PT_ERR_INVALID, "at most " + std::to_string(gpu::kMaxOutputSteps) +
                                         " per-step outputs");
  if (settings.rng == PT_RNG_SOBEL &&
      (settings.max_sample_size == 0 || settings.max_sample_size > 20000))
    return set_error(PT_ERR_INVALID, "max_sample_size must be in [1, 20000]");
  const auto t0 = std::chrono::steady_clock::now();
  // non-empty items, all InitialRays or all InitialIdxAndDir
  std::vector<uint32_t> act;
  int with_idx = -1;
  for (uint32_t i = 0; i < n; ++i) {
    const pt_ray_batch_item &it = items[i];
    if (!it.scene)
      return set_error(PT_ERR_INVALID, "batch item " + std::to_string(i) + ": null scene");
    if (it.n_rays == 0)
      continue;
    if (!it.rays || !it.out)
      return set_error(PT_ERR_INVALID, "batch item " + std::to_string(i) +
                                           ": null rays or output");
    const int has = it.initial_idx ? 1 : 0;
    if (with_idx >= 0 && has != with_idx)
      return set_error(PT_ERR_INVALID,
                       "InitialRays and InitialIdxAndDir items cannot share one batch");
    with_idx = has;
    if (has) {
      const uint32_t nt = uint32_t(it.scene->flat().tris.size());
      for (uint32_t k = 0; k < it.n_rays; ++k)
        if (it.initial_idx[k] >= nt)
          return set_error(PT_ERR_INVALID, "batch item " + std::to_string(i) +
                                               ": triangle index out of range");
    }
    act.push_back(i);
  }
  r.stats = pt_render_stats{};
  if (act.empty()) {
    if (seconds)
      *seconds = 0.0;
    return PT_OK;
  }
  const uint32_t m = uint32_t(act.size());
  HIP_TRY(hipSetDevice(r.device));
  hipStream_t s = r.stream;
  // the items' scenes in the batch arrays (host SAH trees), as pt_render_batch
  std::vector<pt_batch_item> bitems(m);
  for (uint32_t j = 0; j < m; ++j) {
    bitems[j] = pt_batch_item{};
    bitems[j].scene = items[act[j]].scene;
    bitems[j].x_dim = items[act[j]].n_rays;
    bitems[j].y_dim = 1;
  }
  std::vector<gpu::SceneDesc> desc;
  uint64_t n_pixels = 0, total = 0;
  int rc = ensure_batch_scenes(r, bitems.data(), m, spp, false, settings, &desc, &n_pixels,
                               &total);
  if (rc != PT_OK)
    return rc;
  // per-item work divisions, grid blocks, ray and output offsets
  const uint32_t n_img = n_steps ? n_steps : 1;
  std::vector<gpu::MegaBatchItem> mb(m);
  std::vector<uint32_t> block_begin(m + 1, 0);
  std::vector<size_t> final_off(m + 1, 0); // reduced outputs (items with several blocks)
  uint64_t ray_total = 0;
  size_t out_total = 0;
  uint32_t block_size = 0;
  for (uint32_t j = 0; j < m; ++j) {
    const pt_ray_batch_item &it = items[act[j]];
    WorkDivision wd;
    if (!WorkDivision::make(work_division_settings(settings.render_work_division), spp,
                            it.n_rays, 1, &wd) ||
        wd.block_size > 1024)
      return set_error(PT_ERR_INVALID, "invalid render work division settings");
    block_size = wd.block_size;
    gpu::MegaBatchItem &b = mb[j];
    b = gpu::MegaBatchItem{};
    b.node_off = desc[j].node_off;
    b.tri_off = desc[j].tri_off;
    b.shade_off = desc[j].shade_off;
    b.mat_off = desc[j].mat_off;
    b.light_off = desc[j].light_off;
    b.n_lights = desc[j].n_lights;
    b.ray_off = uint32_t(ray_total);
    b.out_off = out_total;
    b.wd = wd;
    const uint64_t blocks = uint64_t(block_begin[j]) + wd.total_num_blocks();
    if (blocks > 0xffffffffull || ray_total + it.n_rays > 0xffffffffull)
      return set_error(PT_ERR_INVALID, "ray batch too large");
    block_begin[j + 1] = uint32_t(blocks);
    ray_total += it.n_rays;
    out_total += size_t(n_img) * wd.num_sample_blocks * it.n_rays * 3;
    final_off[j + 1] =
        final_off[j] + (wd.num_sample_blocks == 1 ? 0 : size_t(n_img) * it.n_rays * 3);
  }
  // rays and triangle indices through pinned staging
  const size_t ray_bytes = size_t(ray_total) * 6 * sizeof(float);
  const size_t idx_bytes = with_idx ? size_t(ray_total) * sizeof(uint32_t) : 0;
  HIP_TRY(r.b_stage_in.ensure(ray_bytes + idx_bytes));
  char *stage = static_cast<char *>(r.b_stage_in.ptr);
  parallel_for(m, [&](size_t j) {
    const pt_ray_batch_item &it = items[act[j]];
    std::memcpy(stage + size_t(mb[j].ray_off) * 6 * sizeof(float), it.rays,
                size_t(it.n_rays) * 6 * sizeof(float));
    if (with_idx)
      std::memcpy(stage + ray_bytes + size_t(mb[j].ray_off) * sizeof(uint32_t), it.initial_idx,
                  size_t(it.n_rays) * sizeof(uint32_t));
  });
  HIP_TRY(r.scratch_rays.upload(reinterpret_cast<const float *>(stage), size_t(ray_total) * 6, s));
  if (with_idx)
    HIP_TRY(r.scratch_idx.upload(reinterpret_cast<const uint32_t *>(stage + ray_bytes),
                                 size_t(ray_total), s));
  HIP_TRY(r.mb_items.upload(mb.data(), m, s));
  HIP_TRY(r.mb_blocks.upload(block_begin.data(), m + 1, s));
  HIP_TRY(r.rgb.ensure(out_total));
  if (final_off[m])
    HIP_TRY(r.rgb2.ensure(final_off[m]));
  HIP_TRY(r.counter.ensure(1));
  HIP_TRY(hipMemsetAsync(r.counter.ptr, 0, sizeof(unsigned long long), s));
  gpu::RngParams rng{};
  rng.type = settings.rng;
  rng.samples_per = spp;
  if (settings.rng == PT_RNG_SOBEL) {
    rc = r.ensure_sobol(settings.max_sample_size, spp);
    if (rc != PT_OK)
      return rc;
    rng.vals = r.sobol_vals.ptr;
    rng.dim_bound = settings.max_sample_size;
    rng.initial_dim_bound = settings.max_sample_size < 64 ? settings.max_sample_size : 64;
  }
  HIP_TRY(hipStreamSynchronize(s)); // the staging buffer is reused below
  gpu::MegaBatchParams bp{};
  bp.items = r.mb_items.ptr;
  bp.block_begin = r.mb_blocks.ptr;
  bp.n_items = m;
  bp.n_steps = n_steps;
  bp.nodes = reinterpret_cast<const float4 *>(r.b_nodes.ptr);
  bp.tris = reinterpret_cast<const float4 *>(r.b_tris.ptr);
  bp.shade = r.b_shade.ptr;
  bp.materials = r.b_mats.ptr;
  bp.light_tris = r.b_light_tris.ptr;
  bp.light_cdf = r.b_light_cdf.ptr;
  bp.settings = integrator_settings(settings);
  bp.rng = rng;
  bp.rays = r.scratch_rays.ptr;
  bp.initial_idx = with_idx ? r.scratch_idx.ptr : nullptr;
  bp.out = r.rgb.ptr;
  bp.ray_counter = r.counter.ptr;
  HIP_TRY(hipEventRecord(r.ev_start, s));
  HIP_TRY(gpu::launch_megakernel_batch(bp, block_size, block_begin[m],
                                       settings.max_blocks_per_launch_gpu, s));
  // items whose rays span several sample blocks: the second reduction pass
  for (uint32_t j = 0; j < m; ++j) {
    const WorkDivision &wd = mb[j].wd;
    if (wd.num_sample_blocks == 1)
      continue;
    const uint32_t nr = items[act[j]].n_rays;
    for (uint32_t k = 0; k < n_img; ++k)
      HIP_TRY(gpu::launch_reduce_sample_blocks(
          r.rgb.ptr + mb[j].out_off + size_t(k) * wd.num_sample_blocks * nr * 3,
          wd.num_sample_blocks, nr, spp, r.rgb2.ptr + final_off[j] + size_t(k) * nr * 3, s));
  }
  HIP_TRY(hipEventRecord(r.ev_stop, s));
  // results: one copy per buffer into pinned staging, then a parallel scatter
  const size_t out_bytes = out_total * sizeof(float), fin_bytes = final_off[m] * sizeof(float);
  HIP_TRY(r.b_stage_out.ensure(out_bytes + fin_bytes + sizeof(unsigned long long)));
  char *ost = static_cast<char *>(r.b_stage_out.ptr);
  HIP_TRY(hipMemcpyAsync(ost, r.rgb.ptr, out_bytes, hipMemcpyDeviceToHost, s));
  if (fin_bytes)
    HIP_TRY(hipMemcpyAsync(ost + out_bytes, r.rgb2.ptr, fin_bytes, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(ost + out_bytes + fin_bytes, r.counter.ptr, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  parallel_for(m, [&](size_t j) {
    const pt_ray_batch_item &it = items[act[j]];
    const size_t bytes = size_t(n_img) * it.n_rays * 3 * sizeof(float);
    const char *src = mb[j].wd.num_sample_blocks == 1
                          ? ost + mb[j].out_off * sizeof(float)
                          : ost + out_bytes + final_off[j] * sizeof(float);
    std::memcpy(it.out, src, bytes);
  });
  unsigned long long rays_traced = 0;
  std::memcpy(&rays_traced, ost + out_bytes + fin_bytes, sizeof(rays_traced));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
  r.stats.kernel_ms = ms;
  r.stats.rays = rays_traced;
  r.stats.samples = ray_total * spp;
  r.stats.mega_launches = 1;
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}

int autotune_batch(pt_renderer &r, const pt_batch_item *items, uint32_t n,
                   uint32_t spp, pt_settings &settings, uint32_t out[5]) {
  const uint32_t pools[] = {1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24};
  const uint32_t blocks[] = {3, 4, 6, 8};
  const uint32_t iters[] = {1, 4};
  const uint32_t pipes[] = {2, 3, 4};
  double best = 1e30;
  uint32_t bp = settings.max_num_samples_per_launch, bb = r.trace_blocks_per_cu,
           bi = r.iters_per_sync, bn = r.n_pipes;
  // Every error return leaves the renderer and the caller's settings as they
  // were before the search (the five tuned values).
  struct Restore {
    pt_renderer &r;
    pt_settings &s;
    uint32_t pool, blocks, iters, pipes;
    bool graphs;
    bool armed = true;
    ~Restore() {
      if (!armed)
        return;
      s.max_num_samples_per_launch = pool;
      r.trace_blocks_per_cu = blocks;
      r.iters_per_sync = iters;
      r.n_pipes = pipes;
      r.use_graphs = graphs;
    }
  } restore{r, settings, bp, bb, bi, bn, r.use_graphs};
  r.use_graphs = true;
  // warm-up (uploads, Sobol table, code objects)
  int rc = render_batch(r, items, n, spp, settings, nullptr);
  if (rc != PT_OK)
    return rc;
  for (uint32_t np : pipes)
    for (uint32_t pool : pools)
      for (uint32_t b : blocks)
        for (uint32_t it : iters) {
          settings.max_num_samples_per_launch = pool;
          r.trace_blocks_per_cu = b;
          r.iters_per_sync = it;
          r.n_pipes = np;
          rc = render_batch(r, items, n, spp, settings, nullptr);
          if (rc != PT_OK)
            return rc;
          double t = r.stats.kernel_ms;
          if (t < best) {
            best = t;
            bp = pool;
            bb = b;
            bi = it;
            bn = np;
          }
        }
  restore.armed = false;
  settings.max_num_samples_per_launch = bp;
  r.trace_blocks_per_cu = bb;
  r.iters_per_sync = bi;
  r.n_pipes = bn;
  r.use_graphs = true; // every candidate ran with graphs
  out[0] = bp;
  out[1] = bb;
  out[2] = bi;
  out[3] = 1;
  out[4] = bn;
  return PT_OK;
}

} // namespace host
} // namespace pt
