// Seeded random scenes for training-data generation: 2-40 boxes / UV spheres
// in [-1, 1]^3 with diffuse albedos uniform in [0.05, 0.95], a triangle budget
// log-uniform in [20, 20000], 1-3 downward-facing quad emitters of radiance
// uniform in [5, 30], a camera on a 3.5-radius sphere looking at the origin and
// a square resolution drawn from {32, 64, 96, 128}.
//
// The reference's generator (src/generate_data/full_scene/scene_generator.cpp:
// 103-287) draws random triangles; the object population here is the one the
// training-data configuration asks for. Same seed -> same scene on any host.
#include "scene_gen.h"

#include <algorithm>
#include <cmath>

namespace pt {
namespace host {
namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * 0x1p-53; }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int integer(int lo, int hi) { // inclusive
    return lo + int(next() % uint64_t(hi - lo + 1));
  }
};

using Tri = std::array<V3f, 3>;

void add_quad(std::vector<Tri> &t, V3f a, V3f b, V3f c, V3f d) {
  t.push_back({a, b, c});
  t.push_back({a, c, d});
}

// Axis-aligned box with outward-facing triangles.
void add_box(std::vector<Tri> &t, V3f lo, V3f hi) {
  float x0 = lo.x, y0 = lo.y, z0 = lo.z, x1 = hi.x, y1 = hi.y, z1 = hi.z;
  add_quad(t, {x0, y1, z0}, {x0, y1, z1}, {x1, y1, z1}, {x1, y1, z0}); // +y
  add_quad(t, {x0, y0, z0}, {x1, y0, z0}, {x1, y0, z1}, {x0, y0, z1}); // -y
  add_quad(t, {x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}); // +z
  add_quad(t, {x0, y0, z0}, {x0, y1, z0}, {x1, y1, z0}, {x1, y0, z0}); // -z
  add_quad(t, {x1, y0, z0}, {x1, y1, z0}, {x1, y1, z1}, {x1, y0, z1}); // +x
  add_quad(t, {x0, y0, z0}, {x0, y0, z1}, {x0, y1, z1}, {x0, y1, z0}); // -x
}

// UV sphere, outward winding: 2 * seg * (rings - 1) triangles.
void add_sphere(std::vector<Tri> &t, V3f c, float r, int rings, int seg) {
  auto p = [&](int i, int j) {
    double th = M_PI * double(i) / rings;
    double ph = 2.0 * M_PI * double(j) / seg;
    return V3f{c.x + float(r * std::sin(th) * std::cos(ph)),
               c.y + float(r * std::cos(th)),
               c.z + float(r * std::sin(th) * std::sin(ph))};
  };
  for (int i = 0; i < rings; ++i) {
    for (int j = 0; j < seg; ++j) {
      V3f a = p(i, j), b = p(i, j + 1), cc = p(i + 1, j + 1), d = p(i + 1, j);
      if (i != 0)
        t.push_back({a, b, d});
      if (i != rings - 1)
        t.push_back({b, cc, d});
    }
  }
}

} // namespace

GeneratedScene generate_scene(uint64_t seed) {
  SplitMix64 rng{seed * 0x2545f4914f6cdd1dull + 0x1234567ull};
  GeneratedScene g;
  const int n_obj = rng.integer(2, 40);
  const double budget =
      std::exp(rng.uniform(std::log(20.0), std::log(20000.0)));
  // objects become spheres while the budget allows (>= 24 triangles each)
  int n_sphere = 0;
  double spare = budget - 12.0 * n_obj;
  if (spare > 0)
    n_sphere = std::min(n_obj, std::max(0, int(rng.uniform() * n_obj + 1)));
  double sphere_budget = n_sphere ? (spare + 12.0 * n_sphere) / n_sphere : 0;

  std::vector<Tri> tris;
  std::vector<uint32_t> tri_mat;
  std::vector<MaterialParams> mats;
  for (int o = 0; o < n_obj; ++o) {
    MaterialParams m;
    m.type = kDiffuse;
    m.diffuse = {float(rng.uniform(0.05, 0.95)), float(rng.uniform(0.05, 0.95)),
                 float(rng.uniform(0.05, 0.95))};
    mats.push_back(m);
    uint32_t mi = uint32_t(mats.size() - 1);
    V3f c = {float(rng.uniform(-0.8, 0.8)), float(rng.uniform(-0.8, 0.8)),
             float(rng.uniform(-0.8, 0.8))};
    size_t before = tris.size();
    if (o < n_sphere) {
      int rings = std::max(3, int(std::lround(std::sqrt(sphere_budget / 4.0))));
      add_sphere(tris, c, float(rng.uniform(0.08, 0.35)), rings, 2 * rings);
    } else {
      V3f h = {float(rng.uniform(0.05, 0.3)), float(rng.uniform(0.05, 0.3)),
               float(rng.uniform(0.05, 0.3))};
      add_box(tris, c - h, c + h);
    }
    tri_mat.resize(tris.size(), mi);
    (void)before;
  }
  const int n_lights = rng.integer(1, 3);
  for (int l = 0; l < n_lights; ++l) {
    MaterialParams m;
    m.type = kDiffuse;
    m.diffuse = {0.5f, 0.5f, 0.5f};
    float le = float(rng.uniform(5.0, 30.0));
    m.emission = {le, le, le};
    mats.push_back(m);
    uint32_t mi = uint32_t(mats.size() - 1);
    float s = float(rng.uniform(0.1, 0.3));
    float y = float(rng.uniform(1.2, 2.0));
    float x = float(rng.uniform(-0.8, 0.8)), z = float(rng.uniform(-0.8, 0.8));
    // facing down (-y)
    add_quad(tris, {x - s, y, z - s}, {x + s, y, z - s}, {x + s, y, z + s},
             {x - s, y, z + s});
    tri_mat.resize(tris.size(), mi);
  }
  add_mesh(g.scene, tris, tri_mat, mats, Affine{}, "generated");
  const int res_choices[4] = {32, 64, 96, 128};
  g.resolution = uint32_t(res_choices[rng.integer(0, 3)]);
  double az = rng.uniform(0.0, 2.0 * M_PI);
  double el = rng.uniform(10.0, 50.0) * M_PI / 180.0;
  V3f pos = {float(3.5 * std::cos(el) * std::cos(az)), float(3.5 * std::sin(el)),
             float(3.5 * std::cos(el) * std::sin(az))};
  g.camera = camera_from_look(-pos, {0.f, 1.f, 0.f}, pos, 50.f, 1.f);
  return g;
}

} // namespace host
} // namespace pt
