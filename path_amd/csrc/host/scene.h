// Host-side scene model, loaders and the flattened world-space form handed to
// the renderer.
#pragma once

#include "../core/shading.h"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace pt {
namespace host {

using V3f = V3<float>;

struct AABB {
  V3f lo{3.4e38f, 3.4e38f, 3.4e38f};
  V3f hi{-3.4e38f, -3.4e38f, -3.4e38f};
  void add(V3f p) {
    lo = {min_(lo.x, p.x), min_(lo.y, p.y), min_(lo.z, p.z)};
    hi = {max_(hi.x, p.x), max_(hi.y, p.y), max_(hi.z, p.z)};
  }
  void add(const AABB &b) {
    add(b.lo);
    add(b.hi);
  }
};

// Affine transform, row-major 3x4: p' = M p + t. (Eigen::Affine3f in the
// reference.)
struct Affine {
  float m[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
  V3f apply(V3f p) const;
  V3f apply_linear(V3f p) const;
  Affine operator*(const Affine &o) const;
  static Affine translate(V3f t);
  static Affine scale(V3f s);
  static Affine rotate(V3f axis, float angle_rad);
};

// Material as produced by the MTL conversion (float parameters). The derived
// constants are computed by make_material<T>() for the precision in use.
// reference: include/scene/tinyobj_material_conversion.h:13-101
struct MaterialParams {
  uint32_t type = kDiffuse;
  V3f diffuse{0, 0, 0};
  V3f specular{0, 0, 0};
  float shininess = 1.f;
  float ior = 1.f;
  float diffuse_weight = 1.f;
  float glossy_weight = 0.f;
  V3f emission{0, 0, 0};

  template <class T> Material<T> to() const {
    return make_material<T>(type, cast3<T>(diffuse), cast3<T>(specular),
                            T(shininess), T(ior), T(diffuse_weight),
                            T(glossy_weight), cast3<T>(emission));
  }
};

struct TriangleData {
  std::array<V3f, 3> normals;
  uint32_t material_idx;
};

struct EmissiveCluster {
  uint32_t material_idx;
  uint32_t start_idx;
  uint32_t end_idx;
  AABB aabb;
};

// Mirrors scene::Scene (reference include/scene/scene.h:29-130): meshes with
// triangle end offsets, instances (object_to_world + mesh index), object-space
// triangles with per-vertex normals and a material index, materials and
// emissive clusters (runs of consecutive triangles with the same emissive
// material).
struct Scene {
  std::vector<uint32_t> mesh_ends;
  std::vector<AABB> mesh_aabbs;
  std::vector<std::string> mesh_paths;
  std::vector<uint32_t> emissive_cluster_ends_per_mesh;
  std::vector<Affine> instance_transforms;
  std::vector<uint32_t> instance_mesh_idx;
  std::vector<std::array<V3f, 3>> triangles;
  std::vector<TriangleData> triangle_data;
  std::vector<MaterialParams> materials;
  std::vector<EmissiveCluster> emissive_clusters;
};

// World-space flattened scene: what the integrators consume.
struct FlatScene {
  std::vector<TriGeom<float>> tris;   // v0, e1, e2
  std::vector<V3f> normals;           // 3 per triangle
  std::vector<uint32_t> tri_mat;      // material per triangle
  std::vector<MaterialParams> materials;
  std::vector<TriGeom<float>> light_tris;
  std::vector<float> light_cdf;
};

struct Camera {
  // film_to_world, row-major 3x4 (p' = M p + t); origin = translation.
  Affine film_to_world;
};

// ---- construction helpers ----
// Flatten instances into world space and build the emissive-triangle CDF.
// reference: include/intersectable_scene/flat_triangle/flat_triangle.h:79-146,
// include/integrate/light_sampler/random_triangle/random_triangle.h:120-161
FlatScene flatten(const Scene &scene);
// Pieces of flatten() for callers that write the arrays in place: the
// triangle and emitter counts, the shading records of every triangle (the
// values make_shade_tri() gives for flatten()'s output) and the emitters with
// their normalised CDF (flat_light_count() entries each).
size_t flat_triangle_count(const Scene &scene);
size_t flat_light_count(const Scene &scene);
void write_shade_tris(const Scene &scene, ShadeTri<float> *out);
void flatten_lights(const Scene &scene, TriGeom<float> *tris, float *cdf);

// reference: src/scene/camera.cpp:4-27
Camera camera_from_look(V3f look, V3f up, V3f pos, float height_angle_deg,
                        float width_height_ratio);

// reference: include/scene/tinyobj_material_conversion.h:13-101. Returns false
// (with a message) for the parameter combinations the reference aborts on.
struct RawMtl {
  std::string name;
  V3f ka{0, 0, 0}, kd{0, 0, 0}, ks{0, 0, 0}, ke{0, 0, 0};
  float ns = 1.f, ni = 1.f;
  int illum = 0;
};
bool convert_material(const RawMtl &m, MaterialParams *out, std::string *err);

// OBJ (+ MTL) load with tinyobjloader-1.1.0 compatible face order and ear-clip
// triangulation (reference: src/scene/scenefile_compat/scenefile_loader.cpp:120-314,
// extern/include/tiny_obj_loader.h:985-1150). Adds one mesh and one instance.
bool load_obj_mesh(Scene &scene, const std::string &path,
                   const Affine &transform, std::string *err);

// CS123 XML scene file (reference: src/scene/scenefile_compat/scenefile_loader.cpp:23-118
// and CS123XmlSceneParser.cpp). Returns false with a message on failure.
bool load_scene_file(const std::string &path, float width_height_ratio,
                     Scene *scene, Camera *camera, std::string *err);

// Add a mesh given in memory (triangles + materials) as one instance.
void add_mesh(Scene &scene, const std::vector<std::array<V3f, 3>> &tris,
              const std::vector<uint32_t> &tri_materials,
              const std::vector<MaterialParams> &mats, const Affine &transform,
              const std::string &name);

} // namespace host
} // namespace pt
