// C ABI: scene, settings and the GPU renderer.
#include "../../include/path_amd.h"

#include "core/camera.h"
#include "core/rng.h"
#include "gpu/kernels.h"
#include "gpu/radix_sort.h"
#include "host/bvh.h"
#include "host/renderer.h"
#include "host/scene_gen.h"
#include "host/scene.h"
#include "host/settings.h"

#include <rocrand/rocrand.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace pt;
using namespace pt::host;

namespace {
thread_local std::string g_error;

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}
} // namespace

namespace pt {
namespace host {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
} // namespace host
} // namespace pt

#define PT_GUARD_BEGIN try {
#define PT_GUARD_END                                                           \
  }                                                                            \
  catch (const std::exception &e) {                                            \
    return fail(PT_ERR_INVALID, std::string("internal error: ") + e.what());   \
  }

extern "C" {

const char *pt_last_error(void) { return g_error.c_str(); }

// ------------------------------- scene ----------------------------------

int pt_scene_create(pt_scene **out) {
  if (!out)
    return fail(PT_ERR_INVALID, "null output pointer");
  *out = new pt_scene();
  return PT_OK;
}

void pt_scene_destroy(pt_scene *scene) { delete scene; }

int pt_scene_load(const char *path, float width_height_ratio, int quiet,
                  pt_scene **out, float film_to_world[12]) {
  PT_GUARD_BEGIN
  if (!path || !out)
    return fail(PT_ERR_INVALID, "null argument");
  auto sc = std::make_unique<pt_scene>();
  Camera cam;
  std::string err;
  if (!load_scene_file(path, width_height_ratio, &sc->scene, &cam, &err))
    return fail(PT_ERR_IO, err);
  (void)quiet;
  if (film_to_world)
    std::memcpy(film_to_world, cam.film_to_world.m, 12 * sizeof(float));
  sc->generation = next_generation();
  *out = sc.release();
  return PT_OK;
  PT_GUARD_END
}

static MaterialParams from_pt(const pt_material &m) {
  MaterialParams p;
  p.type = m.type;
  p.diffuse = {m.diffuse[0], m.diffuse[1], m.diffuse[2]};
  p.specular = {m.specular[0], m.specular[1], m.specular[2]};
  p.shininess = m.shininess;
  p.ior = m.ior;
  p.diffuse_weight = m.diffuse_weight;
  p.glossy_weight = m.glossy_weight;
  p.emission = {m.emission[0], m.emission[1], m.emission[2]};
  return p;
}

static pt_material to_pt(const MaterialParams &p) {
  pt_material m;
  m.type = p.type;
  m.diffuse[0] = p.diffuse.x, m.diffuse[1] = p.diffuse.y, m.diffuse[2] = p.diffuse.z;
  m.specular[0] = p.specular.x, m.specular[1] = p.specular.y, m.specular[2] = p.specular.z;
  m.shininess = p.shininess;
  m.ior = p.ior;
  m.diffuse_weight = p.diffuse_weight;
  m.glossy_weight = p.glossy_weight;
  m.emission[0] = p.emission.x, m.emission[1] = p.emission.y, m.emission[2] = p.emission.z;
  return m;
}

int pt_scene_set_material(pt_scene *scene, uint32_t idx, const pt_material *m) {
  PT_GUARD_BEGIN
  if (!scene || !m)
    return fail(PT_ERR_INVALID, "null argument");
  auto &mats = scene->scene.materials;
  if (idx >= mats.size())
    return fail(PT_ERR_INVALID, "material index " + std::to_string(idx) +
                                    " out of range (" + std::to_string(mats.size()) + ")");
  MaterialParams p = from_pt(*m);
  auto emissive = [](const MaterialParams &q) {
    return q.emission.x != 0.f || q.emission.y != 0.f || q.emission.z != 0.f;
  };
  if (emissive(p) != emissive(mats[idx]))
    return fail(PT_ERR_UNSUPPORTED,
                "a material cannot switch between emissive and non-emissive "
                "(the emitter list is fixed when the mesh is added)");
  if (p.type > 5)
    return fail(PT_ERR_INVALID, "unknown BSDF type");
  mats[idx] = p;
  scene->mat_generation = next_generation();
  return PT_OK;
  PT_GUARD_END
}

static Affine affine_from(const float *t) {
  Affine a;
  if (t)
    std::memcpy(a.m, t, 12 * sizeof(float));
  return a;
}

int pt_scene_add_mesh(pt_scene *scene, const float *verts, const float *normals,
                      const uint32_t *tri_mats, uint32_t n_tris,
                      const pt_material *mats, uint32_t n_mats,
                      const float *transform) {
  PT_GUARD_BEGIN
  if (!scene || (n_tris && (!verts || !tri_mats)) || (n_mats && !mats))
    return fail(PT_ERR_INVALID, "null argument");
  for (uint32_t i = 0; i < n_tris; ++i)
    if (tri_mats[i] >= n_mats)
      return fail(PT_ERR_INVALID, "triangle material index out of range");
  for (uint32_t i = 0; i < n_mats; ++i)
    if (mats[i].type > PT_BSDF_DIELECTRIC_REFRACTIVE)
      return fail(PT_ERR_INVALID, "unknown material type");
  std::vector<std::array<V3f, 3>> tris(n_tris);
  for (uint32_t i = 0; i < n_tris; ++i)
    for (int k = 0; k < 3; ++k)
      tris[i][k] = {verts[9 * i + 3 * k], verts[9 * i + 3 * k + 1],
                    verts[9 * i + 3 * k + 2]};
  std::vector<uint32_t> tm(tri_mats, tri_mats + n_tris);
  std::vector<MaterialParams> mp;
  for (uint32_t i = 0; i < n_mats; ++i)
    mp.push_back(from_pt(mats[i]));
  add_mesh(scene->scene, tris, tm, mp, affine_from(transform), "memory");
  if (normals) {
    uint32_t base = uint32_t(scene->scene.triangles.size()) - n_tris;
    for (uint32_t i = 0; i < n_tris; ++i)
      for (int k = 0; k < 3; ++k)
        scene->scene.triangle_data[base + i].normals[k] =
            normalize(V3f{normals[9 * i + 3 * k], normals[9 * i + 3 * k + 1],
                          normals[9 * i + 3 * k + 2]});
  }
  scene->generation = next_generation();
  return PT_OK;
  PT_GUARD_END
}

int pt_scene_add_obj(pt_scene *scene, const char *path, const float *transform) {
  PT_GUARD_BEGIN
  if (!scene || !path)
    return fail(PT_ERR_INVALID, "null argument");
  std::string err;
  if (!load_obj_mesh(scene->scene, path, affine_from(transform), &err))
    return fail(PT_ERR_IO, err);
  scene->generation = next_generation();
  return PT_OK;
  PT_GUARD_END
}

int pt_scene_counts(const pt_scene *scene, uint32_t *n_tris,
                    uint32_t *n_materials, uint32_t *n_light_tris) {
  PT_GUARD_BEGIN
  if (!scene)
    return fail(PT_ERR_INVALID, "null scene");
  const FlatScene &f = scene->flat();
  if (n_tris)
    *n_tris = uint32_t(f.tris.size());
  if (n_materials)
    *n_materials = uint32_t(f.materials.size());
  if (n_light_tris)
    *n_light_tris = uint32_t(f.light_tris.size());
  return PT_OK;
  PT_GUARD_END
}

int pt_scene_flat(const pt_scene *scene, float *tris, float *normals,
                  uint32_t *tri_mat, pt_material *mats, float *light_tris,
                  float *light_cdf) {
  PT_GUARD_BEGIN
  if (!scene)
    return fail(PT_ERR_INVALID, "null scene");
  const FlatScene &f = scene->flat();
  if (tris)
    std::memcpy(tris, f.tris.data(), f.tris.size() * sizeof(TriGeom<float>));
  if (normals)
    std::memcpy(normals, f.normals.data(), f.normals.size() * sizeof(V3f));
  if (tri_mat)
    std::memcpy(tri_mat, f.tri_mat.data(), f.tri_mat.size() * sizeof(uint32_t));
  if (mats)
    for (size_t i = 0; i < f.materials.size(); ++i)
      mats[i] = to_pt(f.materials[i]);
  if (light_tris)
    std::memcpy(light_tris, f.light_tris.data(),
                f.light_tris.size() * sizeof(TriGeom<float>));
  if (light_cdf)
    std::memcpy(light_cdf, f.light_cdf.data(), f.light_cdf.size() * sizeof(float));
  return PT_OK;
  PT_GUARD_END
}

int pt_camera_from_look(const float look[3], const float up[3],
                        const float pos[3], float height_angle_deg,
                        float width_height_ratio, float film_to_world[12]) {
  if (!look || !up || !pos || !film_to_world)
    return fail(PT_ERR_INVALID, "null argument");
  Camera c = camera_from_look({look[0], look[1], look[2]}, {up[0], up[1], up[2]},
                              {pos[0], pos[1], pos[2]}, height_angle_deg,
                              width_height_ratio);
  std::memcpy(film_to_world, c.film_to_world.m, 12 * sizeof(float));
  return PT_OK;
}

int pt_scene_bvh(const pt_scene *scene, uint32_t *n_nodes, uint32_t *n_refs, float *lo,
                 float *hi, uint32_t *is_leaf, uint32_t *a, uint32_t *b, uint32_t *perm,
                 uint32_t *max_depth) {
  PT_GUARD_BEGIN
  if (!scene || !n_nodes)
    return fail(PT_ERR_INVALID, "null argument");
  const Bvh &bvh = scene->bvh();
  *n_nodes = uint32_t(bvh.nodes.size());
  if (n_refs)
    *n_refs = uint32_t(bvh.perm.size());
  if (max_depth)
    *max_depth = bvh.max_depth;
  for (size_t i = 0; i < bvh.nodes.size(); ++i) {
    const BvhNode &n = bvh.nodes[i];
    if (lo)
      lo[3 * i] = n.lo.x, lo[3 * i + 1] = n.lo.y, lo[3 * i + 2] = n.lo.z;
    if (hi)
      hi[3 * i] = n.hi.x, hi[3 * i + 1] = n.hi.y, hi[3 * i + 2] = n.hi.z;
    if (is_leaf)
      is_leaf[i] = n.is_leaf;
    if (a)
      a[i] = n.a;
    if (b)
      b[i] = n.b;
  }
  if (perm)
    std::memcpy(perm, bvh.perm.data(), bvh.perm.size() * sizeof(uint32_t));
  return PT_OK;
  PT_GUARD_END
}

int pt_scene_set_accel(pt_scene *scene, const pt_settings *settings) {
  PT_GUARD_BEGIN
  if (!scene || !settings)
    return fail(PT_ERR_INVALID, "null argument");
  scene->use_tree_settings(tree_settings(*settings));
  return PT_OK;
  PT_GUARD_END
}

int pt_pack_node(const float *lo, const float *hi, uint32_t n, const int32_t *child,
                 void *node, float *decoded_lo, float *decoded_hi) {
  if (!lo || !hi || !child || !node || n == 0 || n > 4)
    return fail(PT_ERR_INVALID, "null argument or not 1-4 children");
  float clo[4][3], chi[4][3];
  for (uint32_t k = 0; k < n; ++k)
    for (int a = 0; a < 3; ++a) {
      clo[k][a] = lo[3 * k + a];
      chi[k][a] = hi[3 * k + a];
    }
  GpuNode o;
  const bool ok = quantise_gpu_node(clo, chi, n, child, &o);
  std::memcpy(node, &o, sizeof(o));
  // decode exactly as the traversal does: plane = fma(code, step, origin)
  const uint8_t *ql[3] = {o.lo_x, o.lo_y, o.lo_z}, *qh[3] = {o.hi_x, o.hi_y, o.hi_z};
  for (uint32_t k = 0; k < 4; ++k)
    for (int a = 0; a < 3; ++a) {
      if (decoded_lo)
        decoded_lo[3 * k + a] = std::fma(float(ql[a][k]), o.step[a], o.origin[a]);
      if (decoded_hi)
        decoded_hi[3 * k + a] = std::fma(float(qh[a][k]), o.step[a], o.origin[a]);
    }
  return ok ? PT_OK : fail(PT_ERR_INVALID, "quantised node is not conservative");
}

int pt_clip_triangle_bounds(const float tri[9], const float lo[3], const float hi[3],
                            float out_lo[3], float out_hi[3]) {
  if (!tri || !lo || !hi || !out_lo || !out_hi)
    return fail(PT_ERR_INVALID, "null argument");
  const V3<float> v[3] = {{tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]},
                          {tri[6], tri[7], tri[8]}};
  clip_triangle_bounds(v, lo, hi, out_lo, out_hi);
  return PT_OK;
}

// ------------------------------ settings --------------------------------

void pt_settings_default(pt_settings *s) {
  if (s)
    default_settings(s);
}

int pt_settings_load_yaml(const char *path, pt_settings *s) {
  PT_GUARD_BEGIN
  if (!path || !s)
    return fail(PT_ERR_INVALID, "null argument");
  std::string err;
  if (!load_settings_yaml(path, s, &err))
    return fail(PT_ERR_IO, err);
  return PT_OK;
  PT_GUARD_END
}

int pt_settings_to_yaml(const pt_settings *s, char *buf, size_t cap,
                        size_t *needed) {
  PT_GUARD_BEGIN
  if (!s)
    return fail(PT_ERR_INVALID, "null settings");
  std::string y = settings_to_yaml(*s);
  if (needed)
    *needed = y.size();
  if (buf && cap) {
    size_t n = y.size() < cap - 1 ? y.size() : cap - 1;
    std::memcpy(buf, y.data(), n);
    buf[n] = 0;
  }
  return PT_OK;
  PT_GUARD_END
}

int pt_work_division(const pt_work_division_settings *s, uint32_t samples_per,
                     uint32_t x_dim, uint32_t y_dim, uint32_t out[11]) {
  if (!s || !out)
    return fail(PT_ERR_INVALID, "null argument");
  WorkDivision d;
  if (!WorkDivision::make(work_division_settings(*s), samples_per, x_dim, y_dim, &d))
    return fail(PT_ERR_INVALID, "invalid work division settings");
  uint32_t v[11] = {d.block_size,     d.base_samples_per_thread,
                    d.n_threads_per_unit_extra, d.sample_block_size,
                    d.x_block_size,   d.y_block_size,
                    d.num_sample_blocks, d.num_x_blocks,
                    d.num_y_blocks,   d.x_dim,
                    d.y_dim};
  std::memcpy(out, v, sizeof(v));
  return PT_OK;
}

int pt_sobol_direction_numbers(uint32_t dim_bound, uint32_t *out) {
  if (!out || dim_bound == 0 || dim_bound > 20000)
    return fail(PT_ERR_INVALID, "dim_bound must be in [1, 20000]");
  const unsigned int *v = nullptr;
  if (rocrand_get_direction_vectors32(&v, ROCRAND_DIRECTION_VECTORS_32_JOEKUO6) !=
          ROCRAND_STATUS_SUCCESS ||
      !v)
    return fail(PT_ERR_UNSUPPORTED, "Joe-Kuo direction numbers unavailable");
  std::memcpy(out, v, size_t(dim_bound) * 32 * sizeof(uint32_t));
  return PT_OK;
}

// ------------------------------ renderer --------------------------------

int pt_renderer_create(int device, pt_renderer **out) {
  PT_GUARD_BEGIN
  if (!out)
    return fail(PT_ERR_INVALID, "null output pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(PT_ERR_NO_DEVICE, "no GPU device visible");
  if (device < 0 || device >= n)
    return fail(PT_ERR_INVALID, "device index out of range");
  auto r = std::make_unique<pt_renderer>();
  std::string err;
  if (!r->init(device, &err))
    return fail(PT_ERR_HIP, err);
  *out = r.release();
  return PT_OK;
  PT_GUARD_END
}

void pt_renderer_destroy(pt_renderer *r) { delete r; }

int pt_render(pt_renderer *r, const pt_sample_spec *spec, const pt_output *out,
              const pt_scene *scene, uint32_t samples_per,
              const pt_settings *settings, int show_progress, int show_times,
              double *seconds) {
  PT_GUARD_BEGIN
  if (!r || !spec || !out || !scene || !settings)
    return fail(PT_ERR_INVALID, "null argument");
  return r->render(*spec, *out, *scene, samples_per, *settings, show_progress != 0,
                   show_times != 0, seconds);
  PT_GUARD_END
}

int pt_intersect(pt_renderer *r, const pt_scene *scene, const float *rays,
                 uint32_t n, float *t, uint32_t *idx, uint32_t *back) {
  PT_GUARD_BEGIN
  if (!r || !scene || (n && (!rays || !t || !idx || !back)))
    return fail(PT_ERR_INVALID, "null argument");
  return r->intersect(*scene, rays, n, t, idx, back);
  PT_GUARD_END
}

int pt_sobol_table(pt_renderer *r, uint32_t dim_bound, uint32_t count,
                   float *out) {
  PT_GUARD_BEGIN
  if (!r || !out)
    return fail(PT_ERR_INVALID, "null argument");
  return r->sobol_table_to_host(dim_bound, count, out);
  PT_GUARD_END
}

int pt_sort_pairs(pt_renderer *r, const uint64_t *keys, const uint32_t *vals, uint32_t n,
                  uint32_t end_bit, uint64_t *keys_out, uint32_t *vals_out) {
  PT_GUARD_BEGIN
  if (!r || (n && (!keys || !vals || !keys_out || !vals_out)))
    return fail(PT_ERR_INVALID, "null argument");
  if (end_bit < 1 || end_bit > 64)
    return fail(PT_ERR_INVALID, "end_bit must be in [1, 64]");
  if (n == 0)
    return PT_OK;
  hipStream_t s = r->stream;
  DeviceBuf<unsigned long long> k0, k1;
  DeviceBuf<uint32_t> v0, v1;
  DeviceBuf<char> temp;
  const size_t tb = gpu::radix_sort_temp_bytes(n);
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = k0.upload(reinterpret_cast<const unsigned long long *>(keys), n, s);
  if (e == hipSuccess)
    e = v0.upload(vals, n, s);
  if (e == hipSuccess)
    e = k1.ensure(n);
  if (e == hipSuccess)
    e = v1.ensure(n);
  if (e == hipSuccess)
    e = temp.ensure(tb);
  if (e == hipSuccess)
    e = gpu::radix_sort_pairs(k0.ptr, v0.ptr, k1.ptr, v1.ptr, n, end_bit, temp.ptr, tb, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(keys_out, k1.ptr, size_t(n) * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(vals_out, v1.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("sort: ") + hipGetErrorString(e));
  return PT_OK;
  PT_GUARD_END
}

int pt_reduce_samples(pt_renderer *r, const pt_work_division_settings *s,
                      const float *values, uint32_t num_locations, uint32_t samples_per,
                      float *out) {
  PT_GUARD_BEGIN
  if (!r || !s || !values || !out)
    return fail(PT_ERR_INVALID, "null argument");
  if (num_locations == 0 || samples_per == 0)
    return fail(PT_ERR_INVALID, "empty reduction");
  WorkDivision d;
  if (!WorkDivision::make(work_division_settings(*s), samples_per, num_locations, 1, &d))
    return fail(PT_ERR_INVALID, "invalid work division settings");
  if (d.num_sample_blocks != 1)
    return fail(PT_ERR_INVALID, "the reduction needs one sample block per location");
  if (d.block_size > 1024)
    return fail(PT_ERR_INVALID, "block_size above 1024");
  const size_t n = size_t(num_locations) * samples_per;
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = r->scratch_t.upload(values, n, r->stream);
  if (e == hipSuccess)
    e = r->rgb.ensure(size_t(num_locations) * 3);
  if (e == hipSuccess)
    e = gpu::launch_reduce_samples(d, r->scratch_t.ptr, samples_per, r->rgb.ptr, r->stream);
  std::vector<float> rgb(size_t(num_locations) * 3);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rgb.data(), r->rgb.ptr, rgb.size() * sizeof(float),
                       hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("reduce: ") + hipGetErrorString(e));
  for (uint32_t i = 0; i < num_locations; ++i)
    out[i] = rgb[3 * size_t(i)];
  return PT_OK;
  PT_GUARD_END
}

int pt_scene_generate(uint64_t seed, pt_scene **out, float film_to_world[12],
                      uint32_t *resolution) {
  PT_GUARD_BEGIN
  if (!out)
    return fail(PT_ERR_INVALID, "null output pointer");
  GeneratedScene g = generate_scene(seed);
  auto sc = std::make_unique<pt_scene>();
  sc->scene = std::move(g.scene);
  sc->generation = next_generation();
  if (film_to_world)
    std::memcpy(film_to_world, g.camera.film_to_world.m, 12 * sizeof(float));
  if (resolution)
    *resolution = g.resolution;
  *out = sc.release();
  return PT_OK;
  PT_GUARD_END
}

int pt_render_batch(pt_renderer *r, const pt_batch_item *items, uint32_t n,
                    uint32_t samples_per, const pt_settings *settings,
                    double *seconds) {
  PT_GUARD_BEGIN
  if (!r || (n && !items) || !settings)
    return fail(PT_ERR_INVALID, "null argument");
  return render_batch(*r, items, n, samples_per, *settings, seconds);
  PT_GUARD_END
}

int pt_render_rays_batch(pt_renderer *r, const pt_ray_batch_item *items, uint32_t n,
                         uint32_t samples_per, uint32_t n_steps,
                         const pt_settings *settings, double *seconds) {
  PT_GUARD_BEGIN
  if (!r || !settings || (n && !items))
    return fail(PT_ERR_INVALID, "null argument");
  return render_rays_batch(*r, items, n, samples_per, n_steps, *settings, seconds);
  PT_GUARD_END
}

int pt_set_launch_params(pt_renderer *r, uint32_t trace_blocks_per_cu,
                         uint32_t iters_per_sync, int use_graphs) {
  if (!r || trace_blocks_per_cu > 256 || iters_per_sync == 0 || iters_per_sync > 64)
    return fail(PT_ERR_INVALID, "invalid launch parameters");
  r->trace_blocks_per_cu = trace_blocks_per_cu;
  r->iters_per_sync = iters_per_sync;
  r->use_graphs = use_graphs != 0;
  return PT_OK;
}

int pt_set_pipelines(pt_renderer *r, uint32_t n) {
  if (!r || n == 0 || n > uint32_t(kMaxPipes))
    return fail(PT_ERR_INVALID, "pipelines must be in [1, " + std::to_string(kMaxPipes) + "]");
  r->n_pipes = n;
  return PT_OK;
}

// ------------------------------ device buffers ----------------------------

int pt_device_alloc(pt_renderer *r, size_t bytes, void **out) {
  if (!r || !out || bytes == 0)
    return fail(PT_ERR_INVALID, "device buffer needs a renderer, an output and a size");
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = hipMalloc(out, bytes);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("device buffer: ") + hipGetErrorString(e));
  return PT_OK;
}

int pt_device_free(pt_renderer *r, void *ptr) {
  if (!r)
    return fail(PT_ERR_INVALID, "null renderer");
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess && ptr)
    e = hipFree(ptr);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("device buffer free: ") + hipGetErrorString(e));
  return PT_OK;
}

int pt_device_read(pt_renderer *r, const void *src, void *host_dst, size_t bytes) {
  if (!r || !src || !host_dst)
    return fail(PT_ERR_INVALID, "null argument");
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = hipMemcpy(host_dst, src, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("device buffer read: ") + hipGetErrorString(e));
  return PT_OK;
}

int pt_device_write(pt_renderer *r, void *dst, const void *host_src, size_t bytes) {
  if (!r || !dst || !host_src)
    return fail(PT_ERR_INVALID, "null argument");
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = hipMemcpy(dst, host_src, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("device buffer write: ") + hipGetErrorString(e));
  return PT_OK;
}

int pt_l2_loss_gradient(pt_renderer *r, const float *image, const float *target, uint64_t n,
                        float *d_image) {
  if (!r || !image || !target || !d_image)
    return fail(PT_ERR_INVALID, "null argument");
  for (const void *ptr : {static_cast<const void *>(image), static_cast<const void *>(target),
                          static_cast<const void *>(d_image)}) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess || attr.type != hipMemoryTypeDevice ||
        attr.device != r->device) {
      (void)hipGetLastError();
      return fail(PT_ERR_INVALID, "pt_l2_loss_gradient takes device memory of the renderer's GPU");
    }
  }
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = pt::gpu::launch_l2_gradient(image, target, n, d_image, r->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("l2 gradient: ") + hipGetErrorString(e));
  return PT_OK;
}

int pt_set_megakernel_split(pt_renderer *r, int enable) {
  if (!r)
    return fail(PT_ERR_INVALID, "null renderer");
  r->mega_split = enable != 0;
  return PT_OK;
}

int pt_set_bvh_builder(pt_renderer *r, int builder) {
  if (!r || (builder != PT_BVH_HOST_SAH && builder != PT_BVH_GPU_LBVH &&
             builder != PT_BVH_GPU_PLOC))
    return fail(PT_ERR_INVALID, "unknown BVH builder");
  r->gpu_bvh = builder != PT_BVH_HOST_SAH;
  if (r->gpu_bvh)
    r->gpu_builder = builder == PT_BVH_GPU_PLOC ? gpu::kLbvhPloc : gpu::kLbvhRadixTree;
  return PT_OK;
}

int pt_last_bvh_build(const pt_renderer *r, double *ms, uint32_t *levels) {
  if (!r || !ms || !levels)
    return fail(PT_ERR_INVALID, "null argument");
  *ms = r->last_gpu_bvh_ms;
  *levels = r->last_gpu_bvh_levels;
  return PT_OK;
}

int pt_last_batch_phases(const pt_renderer *r, double ms[4]) {
  if (!r || !ms)
    return fail(PT_ERR_INVALID, "null argument");
  for (int k = 0; k < 4; ++k)
    ms[k] = r->last_batch_phase_ms[k];
  return PT_OK;
}

int pt_autotune_batch(pt_renderer *r, const pt_batch_item *items, uint32_t n,
                      uint32_t samples_per, pt_settings *settings,
                      uint32_t out[5]) {
  PT_GUARD_BEGIN
  if (!r || !items || !n || !settings || !out)
    return fail(PT_ERR_INVALID, "null argument");
  return autotune_batch(*r, items, n, samples_per, *settings, out);
  PT_GUARD_END
}

int pt_render_adjoint(pt_renderer *r, const pt_scene *scene, const float film_to_world[12],
                      uint32_t x_dim, uint32_t y_dim, uint32_t samples_per,
                      const pt_settings *settings, const float *d_image,
                      float *grad_albedo, float *grad_emission, double *seconds) {
  PT_GUARD_BEGIN
  if (!r || !scene || !settings)
    return fail(PT_ERR_INVALID, "null argument");
  return render_adjoint(*r, *scene, film_to_world, x_dim, y_dim, samples_per, *settings,
                        d_image, grad_albedo, grad_emission, seconds);
  PT_GUARD_END
}

int pt_last_render_fp64(pt_renderer *r, double *out_rgb) {
  PT_GUARD_BEGIN
  if (!r || !out_rgb)
    return fail(PT_ERR_INVALID, "null argument");
  if (r->last_wf_spp == 0)
    return fail(PT_ERR_INVALID,
                "no Streaming SquareImage render to re-sum (render with kernel_approach "
                "Streaming first)");
  const uint32_t n = r->last_wf_x * r->last_wf_y;
  hipError_t e = hipSetDevice(r->device);
  if (e == hipSuccess)
    e = r->sum64.ensure(size_t(n) * 3);
  if (e == hipSuccess)
    e = gpu::launch_sum_samples_f64(r->wf_samples.ptr, n, r->last_wf_spp, r->sum64.ptr,
                                    r->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_rgb, r->sum64.ptr, size_t(n) * 3 * sizeof(double),
                       hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess)
    return fail(PT_ERR_HIP, std::string("fp64 re-sum: ") + hipGetErrorString(e));
  return PT_OK;
  PT_GUARD_END
}

int pt_set_profiling(pt_renderer *r, int enable) {
  if (!r)
    return fail(PT_ERR_INVALID, "null renderer");
  r->profiling = enable != 0;
  return PT_OK;
}

int pt_last_render_stats(const pt_renderer *r, pt_render_stats *out) {
  if (!r || !out)
    return fail(PT_ERR_INVALID, "null argument");
  *out = r->stats;
  return PT_OK;
}

} // extern "C"
