#include "scene.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace pt {
namespace host {

V3f Affine::apply_linear(V3f p) const {
  return {fma_(m[0][2], p.z, fma_(m[0][1], p.y, m[0][0] * p.x)),
          fma_(m[1][2], p.z, fma_(m[1][1], p.y, m[1][0] * p.x)),
          fma_(m[2][2], p.z, fma_(m[2][1], p.y, m[2][0] * p.x))};
}

V3f Affine::apply(V3f p) const {
  V3f l = apply_linear(p);
  return {l.x + m[0][3], l.y + m[1][3], l.z + m[2][3]};
}

Affine Affine::operator*(const Affine &o) const {
  Affine r;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 4; ++j) {
      double acc = j == 3 ? double(m[i][3]) : 0.0;
      for (int k = 0; k < 3; ++k)
        acc += double(m[i][k]) * double(o.m[k][j]);
      r.m[i][j] = float(acc);
    }
  }
  return r;
}

Affine Affine::translate(V3f t) {
  Affine a;
  a.m[0][3] = t.x;
  a.m[1][3] = t.y;
  a.m[2][3] = t.z;
  return a;
}

Affine Affine::scale(V3f s) {
  Affine a;
  a.m[0][0] = s.x;
  a.m[1][1] = s.y;
  a.m[2][2] = s.z;
  return a;
}

// Eigen::AngleAxis::toRotationMatrix with the axis used as given (the scene
// parser does not normalize it).
Affine Affine::rotate(V3f axis, float angle) {
  double ax = axis.x, ay = axis.y, az = axis.z;
  double s = std::sin(double(angle)), c = std::cos(double(angle));
  double sx = s * ax, sy = s * ay, sz = s * az;
  double cx = (1 - c) * ax, cy = (1 - c) * ay;
  Affine a;
  double t;
  t = cx * ay;
  a.m[0][1] = float(t - sz);
  a.m[1][0] = float(t + sz);
  t = cx * az;
  a.m[0][2] = float(t + sy);
  a.m[2][0] = float(t - sy);
  t = cy * az;
  a.m[1][2] = float(t - sx);
  a.m[2][1] = float(t + sx);
  a.m[0][0] = float(cx * ax + c);
  a.m[1][1] = float(cy * ay + c);
  a.m[2][2] = float((1 - c) * az * az + c);
  return a;
}

static TriGeom<float> tri_geom(V3f a, V3f b, V3f c) {
  return {a, b - a, c - a};
}

// World-space geometry and vertex normals of triangle j under transform tr
// (flatten() and write_shade_tris() both use this, so their values agree).
static inline void flatten_triangle(const Scene &scene, const Affine &tr, uint32_t j,
                                    TriGeom<float> *g, V3f n[3]) {
  const auto &v = scene.triangles[j];
  *g = tri_geom(tr.apply(v[0]), tr.apply(v[1]), tr.apply(v[2]));
  const TriangleData &d = scene.triangle_data[j];
  for (int k = 0; k < 3; ++k)
    n[k] = normalize(tr.apply_linear(d.normals[k]));
}

static inline uint32_t mesh_begin(const Scene &scene, uint32_t mesh) {
  return mesh == 0 ? 0 : scene.mesh_ends[mesh - 1];
}

size_t flat_triangle_count(const Scene &scene) {
  size_t n = 0;
  for (size_t i = 0; i < scene.instance_transforms.size(); ++i) {
    const uint32_t mesh = scene.instance_mesh_idx[i];
    n += scene.mesh_ends[mesh] - mesh_begin(scene, mesh);
  }
  return n;
}

size_t flat_light_count(const Scene &scene) {
  size_t n = 0;
  for (size_t i = 0; i < scene.instance_transforms.size(); ++i) {
    const uint32_t mesh = scene.instance_mesh_idx[i];
    const uint32_t cstart = mesh == 0 ? 0 : scene.emissive_cluster_ends_per_mesh[mesh - 1];
    for (uint32_t c = cstart; c < scene.emissive_cluster_ends_per_mesh[mesh]; ++c)
      n += scene.emissive_clusters[c].end_idx - scene.emissive_clusters[c].start_idx;
  }
  return n;
}

// Emissive triangle CDF weighted by world-space area * emission.sum().
void flatten_lights(const Scene &scene, TriGeom<float> *tris, float *cdf) {
  float cumulative = 0.0f;
  size_t k = 0;
  for (size_t i = 0; i < scene.instance_transforms.size(); ++i) {
    uint32_t mesh = scene.instance_mesh_idx[i];
    const Affine &tr = scene.instance_transforms[i];
    uint32_t cstart = mesh == 0 ? 0 : scene.emissive_cluster_ends_per_mesh[mesh - 1];
    uint32_t cend = scene.emissive_cluster_ends_per_mesh[mesh];
    for (uint32_t c = cstart; c < cend; ++c) {
      const EmissiveCluster &cl = scene.emissive_clusters[c];
      for (uint32_t t = cl.start_idx; t < cl.end_idx; ++t) {
        const auto &v = scene.triangles[t];
        TriGeom<float> g = tri_geom(tr.apply(v[0]), tr.apply(v[1]), tr.apply(v[2]));
        float area = sqrt_(squared_norm(normal_scaled_by_area(g)));
        cumulative += area * sum(scene.materials[cl.material_idx].emission);
        tris[k] = g;
        cdf[k] = cumulative;
        ++k;
      }
    }
  }
  for (size_t j = 0; j < k; ++j)
    cdf[j] /= cumulative;
}

FlatScene flatten(const Scene &scene) {
  FlatScene out;
  out.materials = scene.materials;
  const size_t nt = flat_triangle_count(scene);
  out.tris.resize(nt);
  out.normals.resize(3 * nt);
  out.tri_mat.resize(nt);
  size_t k = 0;
  for (size_t i = 0; i < scene.instance_transforms.size(); ++i) {
    const uint32_t mesh = scene.instance_mesh_idx[i];
    const Affine &tr = scene.instance_transforms[i];
    for (uint32_t j = mesh_begin(scene, mesh); j < scene.mesh_ends[mesh]; ++j, ++k) {
      flatten_triangle(scene, tr, j, &out.tris[k], &out.normals[3 * k]);
      out.tri_mat[k] = scene.triangle_data[j].material_idx;
    }
  }
  const size_t nl = flat_light_count(scene);
  out.light_tris.resize(nl);
  out.light_cdf.resize(nl);
  flatten_lights(scene, out.light_tris.data(), out.light_cdf.data());
  return out;
}

void write_shade_tris(const Scene &scene, ShadeTri<float> *out) {
  size_t k = 0;
  for (size_t i = 0; i < scene.instance_transforms.size(); ++i) {
    const uint32_t mesh = scene.instance_mesh_idx[i];
    const Affine &tr = scene.instance_transforms[i];
    for (uint32_t j = mesh_begin(scene, mesh); j < scene.mesh_ends[mesh]; ++j, ++k) {
      TriGeom<float> g;
      V3f n[3];
      flatten_triangle(scene, tr, j, &g, n);
      out[k] = make_shade_tri(g, n, scene.triangle_data[j].material_idx);
    }
  }
}

Camera camera_from_look(V3f look_f, V3f up_f, V3f pos_f, float height_angle_deg,
                        float width_height_ratio) {
  // f = look, s = f x up, u = s x f; view rows (s, u, -f); film_to_world =
  // inverse(diag(1/tan_w, 1/tan_h, 1) * view). Computed in double.
  auto nrm = [](V3f v) {
    double l = std::sqrt(double(v.x) * v.x + double(v.y) * v.y + double(v.z) * v.z);
    return std::array<double, 3>{v.x / l, v.y / l, v.z / l};
  };
  auto f = nrm(look_f), up = nrm(up_f);
  auto crossd = [](std::array<double, 3> a, std::array<double, 3> b) {
    return std::array<double, 3>{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                                 a[0] * b[1] - a[1] * b[0]};
  };
  auto s = crossd(f, up);
  auto u = crossd(s, f);
  double pos[3] = {pos_f.x, pos_f.y, pos_f.z};
  double half = M_PI * double(height_angle_deg) / 360.0;
  double tan_h = std::tan(half);
  double tan_w = double(width_height_ratio) * tan_h;
  double A[3][3], b[3];
  const std::array<double, 3> rows[3] = {s, u, {-f[0], -f[1], -f[2]}};
  double scl[3] = {1.0 / tan_w, 1.0 / tan_h, 1.0};
  for (int i = 0; i < 3; ++i) {
    double dotp = rows[i][0] * pos[0] + rows[i][1] * pos[1] + rows[i][2] * pos[2];
    for (int j = 0; j < 3; ++j)
      A[i][j] = scl[i] * rows[i][j];
    b[i] = -scl[i] * dotp;
  }
  // inverse of [A | b] = [A^-1 | -A^-1 b]
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  double inv[3][3];
  inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
  inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  Camera cam;
  for (int i = 0; i < 3; ++i) {
    double t = 0;
    for (int j = 0; j < 3; ++j) {
      cam.film_to_world.m[i][j] = float(inv[i][j]);
      t -= inv[i][j] * b[j];
    }
    cam.film_to_world.m[i][3] = float(t);
  }
  return cam;
}

bool convert_material(const RawMtl &m, MaterialParams *out, std::string *err) {
  if (m.illum != 2 && m.illum != 5) {
    *err = "unhandled illum value: " + std::to_string(m.illum) +
           " (only 2 and 5 supported) in material '" + m.name + "'";
    return false;
  }
  const float eps = 1e-8f;
  bool d_nz = squared_norm(m.kd) > eps;
  bool s_nz = squared_norm(m.ks) > eps;
  const float shininess_threshold = 100.f;
  MaterialParams p;
  p.emission = m.ke;
  p.shininess = m.ns;
  p.ior = m.ni;
  if (m.illum == 5) {
    p.type = kDielectricRefractive;
    p.specular = {1.f, 1.f, 1.f};
  } else if (d_nz && s_nz) {
    V3f t = m.kd + m.ks;
    if (!(max_(max_(t.x, t.y), t.z) < 1.f + 1e-4f)) {
      *err = "diffuse + specular exceeds 1 in material '" + m.name + "'";
      return false;
    }
    float dm = sum(m.kd), sm = sum(m.ks);
    float dw = dm / (dm + sm);
    float gw = 1.f - dw;
    p.diffuse = m.kd / V3f{dw, dw, dw};
    p.specular = m.ks / V3f{gw, gw, gw};
    p.diffuse_weight = dw;
    p.glossy_weight = gw;
    p.type = m.ns > shininess_threshold ? kDiffuseMirror : kDiffuseGlossy;
  } else if (d_nz && !s_nz) {
    p.type = kDiffuse;
    p.diffuse = m.kd;
  } else if (m.ns > shininess_threshold) {
    if (d_nz || !s_nz) {
      *err = "diffuse values non-zero or specular values zero for mirror "
             "(unhandled) in material '" + m.name + "'";
      return false;
    }
    p.type = kMirror;
    p.specular = m.ks;
  } else if (s_nz && !d_nz) {
    p.type = kGlossy;
    p.specular = m.ks;
  } else {
    *err = "unhandled material settings in material '" + m.name + "'";
    return false;
  }
  *out = p;
  return true;
}

// Appends one mesh (object-space triangles) plus one instance of it, with the
// emissive-cluster bookkeeping of scenefile_loader.cpp:181-276.
static void finish_mesh(Scene &scene, const std::vector<std::array<V3f, 3>> &tris,
                        const std::vector<std::array<V3f, 3>> &normals,
                        const std::vector<uint32_t> &mats, uint32_t mat_offset,
                        const Affine &transform, const std::string &name) {
  AABB aabb;
  bool in_cluster = false;
  EmissiveCluster cur{};
  auto end_cluster = [&] {
    cur.end_idx = uint32_t(scene.triangles.size());
    scene.emissive_clusters.push_back(cur);
  };
  for (size_t f = 0; f < tris.size(); ++f) {
    uint32_t mat = mats[f] + mat_offset;
    const MaterialParams &mp = scene.materials[mat];
    bool emissive = squared_norm(mp.emission) > 1e-9f;
    for (int k = 0; k < 3; ++k)
      aabb.add(tris[f][k]);
    if (emissive) {
      if (in_cluster && mat != cur.material_idx) {
        end_cluster();
        in_cluster = false;
      }
      if (!in_cluster) {
        cur = EmissiveCluster{};
        cur.material_idx = mat;
        cur.start_idx = uint32_t(scene.triangles.size());
        in_cluster = true;
      }
      for (int k = 0; k < 3; ++k)
        cur.aabb.add(tris[f][k]);
    } else if (in_cluster) {
      end_cluster();
      in_cluster = false;
    }
    scene.triangles.push_back(tris[f]);
    TriangleData d;
    d.normals = normals[f];
    d.material_idx = mat;
    scene.triangle_data.push_back(d);
  }
  if (in_cluster)
    end_cluster();
  uint32_t mesh_idx = uint32_t(scene.mesh_ends.size());
  scene.mesh_ends.push_back(uint32_t(scene.triangles.size()));
  scene.mesh_aabbs.push_back(aabb);
  scene.mesh_paths.push_back(name);
  scene.emissive_cluster_ends_per_mesh.push_back(
      uint32_t(scene.emissive_clusters.size()));
  scene.instance_transforms.push_back(transform);
  scene.instance_mesh_idx.push_back(mesh_idx);
}

static V3f face_normal(const std::array<V3f, 3> &t) {
  return normalize(cross(t[1] - t[0], t[2] - t[0]));
}

void add_mesh(Scene &scene, const std::vector<std::array<V3f, 3>> &tris,
              const std::vector<uint32_t> &tri_materials,
              const std::vector<MaterialParams> &mats, const Affine &transform,
              const std::string &name) {
  uint32_t offset = uint32_t(scene.materials.size());
  for (const auto &m : mats)
    scene.materials.push_back(m);
  std::vector<std::array<V3f, 3>> normals(tris.size());
  for (size_t i = 0; i < tris.size(); ++i) {
    V3f n = face_normal(tris[i]);
    normals[i] = {n, n, n};
  }
  finish_mesh(scene, tris, normals, tri_materials, offset, transform, name);
}

// ------------------------------- OBJ / MTL -------------------------------

static std::string dir_of(const std::string &path) {
  size_t p = path.find_last_of('/');
  return p == std::string::npos ? std::string("") : path.substr(0, p + 1);
}

static bool load_mtl(const std::string &path, std::vector<RawMtl> *out,
                     std::string *err) {
  std::ifstream in(path);
  if (!in) {
    *err = "cannot open material file " + path;
    return false;
  }
  std::string line;
  RawMtl *cur = nullptr;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tok;
    if (!(ss >> tok) || tok[0] == '#')
      continue;
    auto rgb = [&](V3f *v) {
      float a = 0, b = 0, c = 0;
      ss >> a;
      if (!(ss >> b))
        b = c = a;
      else
        ss >> c;
      *v = {a, b, c};
    };
    if (tok == "newmtl") {
      out->emplace_back();
      cur = &out->back();
      ss >> cur->name;
    } else if (!cur) {
      continue;
    } else if (tok == "Ka") {
      rgb(&cur->ka);
    } else if (tok == "Kd") {
      rgb(&cur->kd);
    } else if (tok == "Ks") {
      rgb(&cur->ks);
    } else if (tok == "Ke") {
      rgb(&cur->ke);
    } else if (tok == "Ns") {
      ss >> cur->ns;
    } else if (tok == "Ni") {
      ss >> cur->ni;
    } else if (tok == "illum") {
      ss >> cur->illum;
    }
  }
  return true;
}

struct VIdx {
  int v, vt, vn;
};

static bool fix_index(int idx, int n, int *ret) {
  if (idx > 0) {
    *ret = idx - 1;
    return true;
  }
  if (idx == 0)
    return false;
  *ret = n + idx;
  return true;
}

static bool point_in_tri(const float *vx, const float *vy, float tx, float ty) {
  bool c = false;
  for (int i = 0, j = 2; i < 3; j = i++) {
    if (((vy[i] > ty) != (vy[j] > ty)) &&
        (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]))
      c = !c;
  }
  return c;
}

// Ear clipping exactly in the order of tiny_obj_loader.h:1000-1150: projection
// axes picked from the first non-degenerate corner, ears searched from
// guess_vert forward, the middle vertex of each ear removed.
static void triangulate(const std::vector<VIdx> &face, const std::vector<float> &v,
                        std::vector<std::array<VIdx, 3>> *out) {
  size_t n = face.size();
  if (n == 3) {
    out->push_back({face[0], face[1], face[2]});
    return;
  }
  size_t axes[2] = {1, 2};
  for (size_t k = 0; k < n; ++k) {
    const float *a = &v[3 * face[k].v], *b = &v[3 * face[(k + 1) % n].v],
                *c = &v[3 * face[(k + 2) % n].v];
    float e0x = b[0] - a[0], e0y = b[1] - a[1], e0z = b[2] - a[2];
    float e1x = c[0] - b[0], e1y = c[1] - b[1], e1z = c[2] - b[2];
    float cx = std::fabs(e0y * e1z - e0z * e1y);
    float cy = std::fabs(e0z * e1x - e0x * e1z);
    float cz = std::fabs(e0x * e1y - e0y * e1x);
    const float eps = 0.0001f;
    if (cx > eps || cy > eps || cz > eps) {
      if (!(cx > cy && cx > cz)) {
        axes[0] = 0;
        if (cz > cx && cz > cy)
          axes[1] = 1;
      }
      break;
    }
  }
  float area = 0;
  for (size_t k = 0; k < n; ++k) {
    const float *a = &v[3 * face[k].v], *b = &v[3 * face[(k + 1) % n].v];
    area += (a[axes[0]] * b[axes[1]] - a[axes[1]] * b[axes[0]]) * 0.5f;
  }
  std::vector<VIdx> rem = face;
  size_t guess = 0;
  int max_rounds = 10;
  while (rem.size() > 3 && max_rounds > 0) {
    size_t np = rem.size();
    if (guess >= np) {
      max_rounds -= 1;
      guess -= np;
    }
    VIdx ind[3];
    float vx[3], vy[3];
    for (int k = 0; k < 3; ++k) {
      ind[k] = rem[(guess + k) % np];
      vx[k] = v[3 * ind[k].v + axes[0]];
      vy[k] = v[3 * ind[k].v + axes[1]];
    }
    float e0x = vx[1] - vx[0], e0y = vy[1] - vy[0];
    float e1x = vx[2] - vx[1], e1y = vy[2] - vy[1];
    float crs = e0x * e1y - e0y * e1x;
    if (crs * area < 0.0f) {
      guess += 1;
      continue;
    }
    bool overlap = false;
    for (size_t o = 3; o < np; ++o) {
      const float *p = &v[3 * rem[(guess + o) % np].v];
      if (point_in_tri(vx, vy, p[axes[0]], p[axes[1]])) {
        overlap = true;
        break;
      }
    }
    if (overlap) {
      guess += 1;
      continue;
    }
    out->push_back({ind[0], ind[1], ind[2]});
    size_t r = (guess + 1) % np;
    while (r + 1 < np) {
      rem[r] = rem[r + 1];
      r += 1;
    }
    rem.pop_back();
  }
  if (rem.size() == 3)
    out->push_back({rem[0], rem[1], rem[2]});
}

bool load_obj_mesh(Scene &scene, const std::string &path, const Affine &transform,
                   std::string *err) {
  std::ifstream in(path);
  if (!in) {
    *err = "cannot open mesh file " + path;
    return false;
  }
  std::vector<float> verts, norms;
  std::vector<RawMtl> raw;
  std::map<std::string, int> mat_ids;
  int cur_mat = -1;
  std::vector<std::array<V3f, 3>> tris, tri_normals;
  std::vector<uint32_t> tri_mats;
  std::string line;
  while (std::getline(in, line)) {
    for (char &c : line)
      if (c == '\t' || c == '\r')
        c = ' ';
    std::istringstream ss(line);
    std::string tok;
    if (!(ss >> tok) || tok[0] == '#')
      continue;
    if (tok == "v") {
      float x = 0, y = 0, z = 0;
      ss >> x >> y >> z;
      verts.insert(verts.end(), {x, y, z});
    } else if (tok == "vn") {
      float x = 0, y = 0, z = 0;
      ss >> x >> y >> z;
      norms.insert(norms.end(), {x, y, z});
    } else if (tok == "mtllib") {
      std::string f;
      ss >> f;
      size_t before = raw.size();
      if (!load_mtl(dir_of(path) + f, &raw, err))
        return false;
      for (size_t i = before; i < raw.size(); ++i)
        mat_ids[raw[i].name] = int(i);
    } else if (tok == "usemtl") {
      std::string name;
      ss >> name;
      auto it = mat_ids.find(name);
      cur_mat = it == mat_ids.end() ? -1 : it->second;
    } else if (tok == "f") {
      std::vector<VIdx> face;
      std::string vs;
      int nv = int(verts.size() / 3), nn = int(norms.size() / 3);
      while (ss >> vs) {
        VIdx vi{-1, -1, -1};
        int a = 0, b = 0, c = 0;
        bool hb = false, hc = false;
        size_t p1 = vs.find('/');
        a = std::atoi(vs.substr(0, p1).c_str());
        if (p1 != std::string::npos) {
          size_t p2 = vs.find('/', p1 + 1);
          std::string sb = vs.substr(p1 + 1, p2 == std::string::npos ? std::string::npos : p2 - p1 - 1);
          if (!sb.empty()) {
            b = std::atoi(sb.c_str());
            hb = true;
          }
          if (p2 != std::string::npos) {
            c = std::atoi(vs.substr(p2 + 1).c_str());
            hc = true;
          }
        }
        if (!fix_index(a, nv, &vi.v) || vi.v < 0 || vi.v >= nv) {
          *err = "bad vertex index in " + path;
          return false;
        }
        (void)hb;
        (void)b;
        if (hc && !fix_index(c, nn, &vi.vn))
          vi.vn = -1;
        face.push_back(vi);
      }
      if (face.size() < 3)
        continue;
      if (cur_mat < 0) {
        *err = "face without a known material in " + path;
        return false;
      }
      std::vector<std::array<VIdx, 3>> out;
      triangulate(face, verts, &out);
      for (auto &t : out) {
        std::array<V3f, 3> tv, tn;
        bool has_n[3];
        for (int k = 0; k < 3; ++k) {
          const float *p = &verts[3 * t[k].v];
          tv[k] = {p[0], p[1], p[2]};
          has_n[k] = t[k].vn >= 0 && t[k].vn < nn;
          if (has_n[k]) {
            const float *q = &norms[3 * t[k].vn];
            tn[k] = normalize(V3f{q[0], q[1], q[2]});
          }
        }
        V3f fn = face_normal(tv);
        for (int k = 0; k < 3; ++k)
          if (!has_n[k])
            tn[k] = fn;
        tris.push_back(tv);
        tri_normals.push_back(tn);
        tri_mats.push_back(uint32_t(cur_mat));
      }
    }
  }
  uint32_t offset = uint32_t(scene.materials.size());
  for (const RawMtl &r : raw) {
    MaterialParams p;
    if (!convert_material(r, &p, err))
      return false;
    scene.materials.push_back(p);
  }
  finish_mesh(scene, tris, tri_normals, tri_mats, offset, transform, path);
  return true;
}

// ------------------------------- XML scene -------------------------------

namespace {
struct XmlNode {
  std::string tag;
  std::map<std::string, std::string> attrs;
  std::vector<std::unique_ptr<XmlNode>> children;
  const XmlNode *child(const std::string &t) const {
    for (auto &c : children)
      if (c->tag == t)
        return c.get();
    return nullptr;
  }
  bool has(const std::string &a) const { return attrs.count(a) != 0; }
  float f(const std::string &a, float dflt = 0.f) const {
    auto it = attrs.find(a);
    return it == attrs.end() ? dflt : float(std::atof(it->second.c_str()));
  }
};

struct XmlParser {
  const std::string &s;
  size_t i = 0;
  std::string err;
  explicit XmlParser(const std::string &src) : s(src) {}
  void skip_ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i]))
      ++i;
  }
  bool skip_misc() {
    for (;;) {
      skip_ws();
      if (s.compare(i, 4, "<!--") == 0) {
        size_t e = s.find("-->", i);
        if (e == std::string::npos)
          return false;
        i = e + 3;
      } else if (s.compare(i, 2, "<?") == 0 || s.compare(i, 2, "<!") == 0) {
        size_t e = s.find('>', i);
        if (e == std::string::npos)
          return false;
        i = e + 1;
      } else if (i < s.size() && s[i] != '<') {
        ++i; // text content is ignored
      } else {
        return true;
      }
    }
  }
  std::unique_ptr<XmlNode> parse_element() {
    if (!skip_misc() || i >= s.size() || s[i] != '<') {
      err = "expected element";
      return nullptr;
    }
    ++i;
    auto node = std::make_unique<XmlNode>();
    while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != '>' &&
           s[i] != '/')
      node->tag += s[i++];
    for (;;) {
      skip_ws();
      if (i >= s.size()) {
        err = "unterminated tag <" + node->tag + ">";
        return nullptr;
      }
      if (s[i] == '/') {
        i += 2; // "/>"
        return node;
      }
      if (s[i] == '>') {
        ++i;
        break;
      }
      std::string name;
      while (i < s.size() && s[i] != '=' && !std::isspace((unsigned char)s[i]))
        name += s[i++];
      skip_ws();
      if (i >= s.size() || s[i] != '=') {
        err = "bad attribute in <" + node->tag + ">";
        return nullptr;
      }
      ++i;
      skip_ws();
      char q = s[i++];
      size_t e = s.find(q, i);
      if (e == std::string::npos) {
        err = "unterminated attribute";
        return nullptr;
      }
      node->attrs[name] = s.substr(i, e - i);
      i = e + 1;
    }
    for (;;) {
      if (!skip_misc()) {
        err = "unterminated element <" + node->tag + ">";
        return nullptr;
      }
      if (s.compare(i, 2, "</") == 0) {
        size_t e = s.find('>', i);
        if (e == std::string::npos) {
          err = "unterminated closing tag";
          return nullptr;
        }
        i = e + 1;
        return node;
      }
      auto c = parse_element();
      if (!c)
        return nullptr;
      node->children.push_back(std::move(c));
    }
  }
};

struct XmlLoader {
  Scene *scene;
  std::string base_dir;
  std::map<std::string, const XmlNode *> trees;
  std::map<std::string, std::pair<uint32_t, AABB>> loaded;
  std::string *err;

  bool add_mesh_file(const std::string &file, const Affine &transform) {
    std::string path = base_dir + file;
    auto it = loaded.find(path);
    if (it != loaded.end()) { // instance of an already loaded mesh
      scene->instance_transforms.push_back(transform);
      scene->instance_mesh_idx.push_back(it->second.first);
      return true;
    }
    if (!load_obj_mesh(*scene, path, transform, err))
      return false;
    loaded[path] = {uint32_t(scene->mesh_ends.size() - 1), scene->mesh_aabbs.back()};
    return true;
  }

  // A <matrix> of four child elements (attributes a-d or v1-v4): the k-th
  // element fills column k (CS123XmlSceneParser.cpp:236-262), and it composes
  // as a full 4x4 product on the right (scenefile_loader.cpp:86-87).
  bool compose_matrix(const XmlNode &e, Affine *t) {
    static const char *const kNames[2][4] = {{"a", "b", "c", "d"}, {"v1", "v2", "v3", "v4"}};
    float mat[4][4];
    int col = 0;
    for (auto &c : e.children) {
      if (col == 4)
        break;
      int which = -1;
      for (int w = 0; w < 2 && which < 0; ++w)
        if (c->has(kNames[w][0]) && c->has(kNames[w][1]) && c->has(kNames[w][2]) &&
            c->has(kNames[w][3]))
          which = w;
      if (which < 0) {
        *err = "<matrix> element needs attributes a, b, c, d (or v1-v4)";
        return false;
      }
      for (int i = 0; i < 4; ++i)
        mat[i][col] = c->f(kNames[which][i]);
      ++col;
    }
    if (col != 4) {
      *err = "<matrix> needs four elements";
      return false;
    }
    Affine r;
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 4; ++k)
        r.m[i][k] = float(double(t->m[i][0]) * mat[0][k] + double(t->m[i][1]) * mat[1][k] +
                          double(t->m[i][2]) * mat[2][k] + double(t->m[i][3]) * mat[3][k]);
    *t = r;
    return true;
  }

  // One scene-graph node per <transblock>: its transformations in order,
  // then its primitives, then its subtrees and master references in order
  // (scenefile_loader.cpp:70-102), so triangles are numbered as there.
  bool parse_tree(const XmlNode &obj, const Affine &parent, int depth) {
    if (depth > 64) {
      *err = "scene graph too deep (cycle?)";
      return false;
    }
    for (auto &tb : obj.children) {
      if (tb->tag != "transblock")
        continue;
      Affine t = parent;
      for (auto &e : tb->children) {
        if (e->tag == "translate") {
          t = t * Affine::translate({e->f("x"), e->f("y"), e->f("z")});
        } else if (e->tag == "scale") {
          t = t * Affine::scale({e->f("x"), e->f("y"), e->f("z")});
        } else if (e->tag == "rotate") {
          float ang = float(double(e->f("angle")) * M_PI / 180.0);
          t = t * Affine::rotate({e->f("x"), e->f("y"), e->f("z")}, ang);
        } else if (e->tag == "matrix") {
          if (!compose_matrix(*e, &t))
            return false;
        }
      }
      for (auto &e : tb->children) {
        if (e->tag != "object" || e->attrs["type"] != "primitive")
          continue;
        auto name = e->attrs["name"];
        if (name != "mesh") {
          *err = "only mesh primitives are supported (got '" + name + "')";
          return false;
        }
        std::string file = e->has("meshfile") ? e->attrs["meshfile"] : e->attrs["filename"];
        if (!add_mesh_file(file, t))
          return false;
      }
      for (auto &e : tb->children) {
        if (e->tag != "object")
          continue;
        auto ty = e->attrs["type"];
        if (ty == "tree") {
          if (!parse_tree(*e, t, depth + 1))
            return false;
        } else if (ty == "master") {
          auto it = trees.find(e->attrs["name"]);
          if (it == trees.end()) {
            *err = "unknown master object '" + e->attrs["name"] + "'";
            return false;
          }
          if (!parse_tree(*it->second, t, depth + 1))
            return false;
        }
      }
    }
    return true;
  }
};
} // namespace

bool load_scene_file(const std::string &path, float width_height_ratio,
                     Scene *scene, Camera *camera, std::string *err) {
  std::ifstream in(path);
  if (!in) {
    *err = "cannot open scene file " + path;
    return false;
  }
  std::stringstream buf;
  buf << in.rdbuf();
  std::string src = buf.str();
  XmlParser p(src);
  auto root = p.parse_element();
  if (!root) {
    *err = "XML parse error in " + path + ": " + p.err;
    return false;
  }
  if (root->tag != "scenefile") {
    *err = "missing <scenefile> root in " + path;
    return false;
  }
  V3f pos{5, 5, 5}, look{-1, -1, -1}, up{0, 1, 0};
  float height_angle = 45.f;
  bool focus = false;
  if (const XmlNode *cam = root->child("cameradata")) {
    for (auto &e : cam->children) {
      if (e->tag == "pos")
        pos = {e->f("x"), e->f("y"), e->f("z")};
      else if (e->tag == "look")
        look = {e->f("x"), e->f("y"), e->f("z")};
      else if (e->tag == "focus") {
        look = {e->f("x"), e->f("y"), e->f("z")};
        focus = true;
      } else if (e->tag == "up")
        up = {e->f("x"), e->f("y"), e->f("z")};
      else if (e->tag == "heightangle")
        height_angle = e->f("v");
    }
  }
  if (focus)
    look = look - pos;
  for (auto &e : root->children) {
    if (e->tag == "lightdata") {
      *err = "point/directional lights not currently supported";
      return false;
    }
  }
  XmlLoader ld;
  ld.scene = scene;
  ld.base_dir = dir_of(path);
  ld.err = err;
  const XmlNode *root_tree = nullptr;
  for (auto &e : root->children) {
    if (e->tag == "object" && e->attrs["type"] == "tree") {
      ld.trees[e->attrs["name"]] = e.get();
      if (e->attrs["name"] == "root")
        root_tree = e.get();
    }
  }
  if (!root_tree) {
    *err = "scene file has no root object";
    return false;
  }
  if (!ld.parse_tree(*root_tree, Affine{}, 0))
    return false;
  if (scene->mesh_ends.empty()) {
    *err = "scene file contains no meshes";
    return false;
  }
  *camera = camera_from_look(look, up, pos, height_angle, width_height_ratio);
  return true;
}

} // namespace host
} // namespace pt
