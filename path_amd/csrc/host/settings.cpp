// Settings defaults and the YAML form of render::Settings.
// reference: include/render/settings.h:27-70, src/render/config_io.cpp:14-34,
// include/lib/serialize_tagged_union.h:11-27 ({type: Enum, value: {...}}),
// configs/benchmark.yaml:1-55 (schema example).
#include "settings.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <vector>

namespace pt {
namespace host {

void default_work_division(pt_work_division_settings *w) {
  w->block_size = 256;
  w->target_x_block_size = 32;
  w->force_target_samples = 0;
  w->forced_target_samples_per_thread = 8;
  w->base_num_threads = 16384;
  w->samples_per_thread_scaling_power = 0.5f;
  w->max_samples_per_thread = 32;
}

void default_settings(pt_settings *s) {
  std::memset(s, 0, sizeof(*s));
  s->kernel_approach = PT_MEGA_KERNEL;
  s->max_blocks_per_launch_gpu = 4096;
  s->max_blocks_per_launch_cpu = 128;
  default_work_division(&s->render_work_division);
  default_work_division(&s->reduce_work_division);
  s->reduce_work_division.target_x_block_size = 256;
  default_work_division(&s->init_samples_division);
  s->max_num_samples_per_launch = 4194304;
  s->min_num_init_blocks = 16;
  s->to_bulk_max_size = 2097152;
  s->accel = PT_ACCEL_SBVH;
  s->overlap_threshold = 1e-5f;
  s->use_spatial_splits = 0;
  s->traversal_per_intersect_cost = 0.5f;
  s->target_objects = 4;
  s->print_stats = 0;
  s->light_sampler = PT_LIGHT_RANDOM_TRIANGLE;
  s->binary_search_threshold = 0xffffffffu / 4;
  s->dir_sampler = PT_DIR_BSDF;
  s->term_prob = PT_TERM_MULTIPLIER_FUNC;
  s->constant_prob = 0.5f;
  s->n_iters = 1;
  s->multiplier_func_exp = 50.0f;
  s->multiplier_func_min_prob = 0.1f;
  s->normalize_min_prob = 0.1f;
  s->rng = PT_RNG_SOBEL;
  s->max_sample_size = 256;
  s->back_cull_emission = 1;
}

IntegratorSettings integrator_settings(const pt_settings &s) {
  IntegratorSettings o{};
  o.light_sampler = s.light_sampler;
  o.dir_sampler = s.dir_sampler;
  o.term_prob = s.term_prob;
  o.rng = s.rng;
  o.back_cull_emission = s.back_cull_emission;
  o.n_iters = s.n_iters;
  o.constant_prob = s.constant_prob;
  o.mf_exp = s.multiplier_func_exp;
  o.mf_min_prob = s.multiplier_func_min_prob;
  o.normalize_min_prob = s.normalize_min_prob;
  return o;
}

WorkDivisionSettings work_division_settings(const pt_work_division_settings &w) {
  WorkDivisionSettings o;
  o.block_size = w.block_size;
  o.target_x_block_size = w.target_x_block_size;
  o.force_target_samples = w.force_target_samples;
  o.forced_target_samples_per_thread = w.forced_target_samples_per_thread;
  o.base_num_threads = w.base_num_threads;
  o.samples_per_thread_scaling_power = w.samples_per_thread_scaling_power;
  o.max_samples_per_thread = w.max_samples_per_thread;
  return o;
}

// ---------------------------------------------------------------------------
// Minimal block-style YAML reader: nested mappings by indentation, scalars,
// `{}` empty maps, comments. Enough for the reference's cereal output.
// ---------------------------------------------------------------------------
namespace {
struct YNode {
  std::string scalar;
  std::vector<std::pair<std::string, std::unique_ptr<YNode>>> map;
  const YNode *get(const std::string &k) const {
    for (auto &kv : map)
      if (kv.first == k)
        return kv.second.get();
    return nullptr;
  }
};

struct Line {
  int indent;
  std::string key, value;
};

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos)
    return "";
  size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

bool parse_lines(std::istream &in, std::vector<Line> *out, std::string *err) {
  std::string raw;
  int ln = 0;
  while (std::getline(in, raw)) {
    ++ln;
    // strip comments that are not inside quotes
    bool q = false;
    for (size_t i = 0; i < raw.size(); ++i) {
      if (raw[i] == '"' || raw[i] == '\'')
        q = !q;
      if (raw[i] == '#' && !q) {
        raw = raw.substr(0, i);
        break;
      }
    }
    if (trim(raw).empty() || trim(raw) == "---")
      continue;
    int indent = 0;
    while (indent < int(raw.size()) && raw[indent] == ' ')
      ++indent;
    std::string body = trim(raw);
    size_t c = body.find(':');
    if (c == std::string::npos) {
      *err = "YAML line " + std::to_string(ln) + ": expected 'key: value'";
      return false;
    }
    Line l;
    l.indent = indent;
    l.key = trim(body.substr(0, c));
    l.value = trim(body.substr(c + 1));
    if (l.value.size() >= 2 && (l.value.front() == '"' || l.value.front() == '\''))
      l.value = l.value.substr(1, l.value.size() - 2);
    out->push_back(l);
  }
  return true;
}

size_t build(const std::vector<Line> &lines, size_t i, int indent, YNode *node) {
  while (i < lines.size() && lines[i].indent >= indent) {
    if (lines[i].indent > indent) { // stray deeper line: attach to previous
      ++i;
      continue;
    }
    auto child = std::make_unique<YNode>();
    const Line &l = lines[i];
    ++i;
    if (l.value.empty()) {
      if (i < lines.size() && lines[i].indent > indent)
        i = build(lines, i, lines[i].indent, child.get());
    } else if (l.value != "{}") {
      child->scalar = l.value;
    }
    node->map.emplace_back(l.key, std::move(child));
  }
  return i;
}

struct Reader {
  std::string *err;
  bool ok = true;
  void fail(const std::string &m) {
    if (ok)
      *err = m;
    ok = false;
  }
  void u32(const YNode *n, const char *k, uint32_t *v) {
    if (!n)
      return;
    if (const YNode *c = n->get(k))
      *v = uint32_t(std::strtoul(c->scalar.c_str(), nullptr, 10));
  }
  void f32(const YNode *n, const char *k, float *v) {
    if (!n)
      return;
    if (const YNode *c = n->get(k))
      *v = float(std::strtod(c->scalar.c_str(), nullptr));
  }
  void boolean(const YNode *n, const char *k, uint32_t *v) {
    if (!n)
      return;
    if (const YNode *c = n->get(k))
      *v = (c->scalar == "true" || c->scalar == "1") ? 1u : 0u;
  }
  // {type: Name, value: {...}} -> enum index
  const YNode *tagged(const YNode *n, const char *k, const char *const *names,
                      uint32_t count, uint32_t *out) {
    if (!n)
      return nullptr;
    const YNode *c = n->get(k);
    if (!c)
      return nullptr;
    const YNode *t = c->get("type");
    if (!t) {
      fail(std::string("missing 'type' under '") + k + "'");
      return nullptr;
    }
    for (uint32_t i = 0; i < count; ++i) {
      if (t->scalar == names[i]) {
        *out = i;
        return c->get("value");
      }
    }
    fail("unknown " + std::string(k) + " type '" + t->scalar + "'");
    return nullptr;
  }
  void work_division(const YNode *n, const char *k, pt_work_division_settings *w) {
    if (!n)
      return;
    const YNode *c = n->get(k);
    if (!c)
      return;
    u32(c, "block_size", &w->block_size);
    u32(c, "target_x_block_size", &w->target_x_block_size);
    boolean(c, "force_target_samples", &w->force_target_samples);
    u32(c, "forced_target_samples_per_thread", &w->forced_target_samples_per_thread);
    u32(c, "base_num_threads", &w->base_num_threads);
    f32(c, "samples_per_thread_scaling_power", &w->samples_per_thread_scaling_power);
    u32(c, "max_samples_per_thread", &w->max_samples_per_thread);
  }
  void accel(const YNode *n, pt_settings *s) {
    static const char *names[] = {"LoopAll", "NaivePartitionBVH", "SBVH",
                                  "DirectionGrid"};
    const YNode *v = tagged(n, "accel", names, 4, &s->accel);
    if (v && s->accel == PT_ACCEL_SBVH) {
      f32(v, "overlap_threshold", &s->overlap_threshold);
      boolean(v, "use_spatial_splits", &s->use_spatial_splits);
      if (const YNode *b = v->get("bvh_settings")) {
        f32(b, "traversal_per_intersect_cost", &s->traversal_per_intersect_cost);
        u32(b, "target_objects", &s->target_objects);
        boolean(b, "print_stats", &s->print_stats);
      }
    }
  }
};

const char *const kApproach[] = {"MegaKernel", "Streaming"};
const char *const kLight[] = {"NoLightSampling", "RandomTriangle"};
const char *const kDir[] = {"Uniform", "BSDF"};
const char *const kTerm[] = {"Constant", "NIters", "MultiplierFunc", "Normalize"};
const char *const kRngNames[] = {"Uniform", "Sobel"};
const char *const kAccel[] = {"LoopAll", "NaivePartitionBVH", "SBVH", "DirectionGrid"};
} // namespace

bool load_settings_yaml(const std::string &path, pt_settings *s,
                        std::string *err) {
  std::ifstream in(path);
  if (!in) {
    *err = "cannot open config file " + path;
    return false;
  }
  std::vector<Line> lines;
  if (!parse_lines(in, &lines, err))
    return false;
  YNode root;
  if (!lines.empty())
    build(lines, 0, lines[0].indent, &root);
  const YNode *st = root.get("settings");
  if (!st) {
    *err = "config has no top-level 'settings'";
    return false;
  }
  pt_settings o;
  default_settings(&o);
  Reader r{err};
  const YNode *ka = r.tagged(st, "kernel_approach", kApproach, 2, &o.kernel_approach);
  if (ka) {
    const YNode *cs = ka->get("computation_settings");
    if (o.kernel_approach == PT_MEGA_KERNEL) {
      r.u32(cs, "max_blocks_per_launch_gpu", &o.max_blocks_per_launch_gpu);
      r.u32(cs, "max_blocks_per_launch_cpu", &o.max_blocks_per_launch_cpu);
      r.work_division(cs, "render_work_division", &o.render_work_division);
      r.work_division(cs, "reduce_work_division", &o.reduce_work_division);
      r.accel(ka->get("individually_intersectable_settings"), &o);
    } else {
      r.work_division(cs, "init_samples_division", &o.init_samples_division);
      r.u32(cs, "max_num_samples_per_launch", &o.max_num_samples_per_launch);
      r.u32(cs, "min_num_init_blocks", &o.min_num_init_blocks);
      static const char *bulk[] = {"IndividualToBulk"};
      uint32_t b = 0;
      const YNode *bv = r.tagged(ka, "accel", bulk, 1, &b);
      if (bv) {
        if (const YNode *tb = bv->get("to_bulk_settings"))
          r.u32(tb, "max_size", &o.to_bulk_max_size);
        r.accel(bv->get("individual_settings"), &o);
      }
    }
  }
  const YNode *ls = r.tagged(st, "light_sampler", kLight, 2, &o.light_sampler);
  if (ls)
    r.u32(ls, "binary_search_threshold", &o.binary_search_threshold);
  r.tagged(st, "dir_sampler", kDir, 2, &o.dir_sampler);
  const YNode *tp = r.tagged(st, "term_prob", kTerm, 4, &o.term_prob);
  if (tp) {
    if (o.term_prob == PT_TERM_CONSTANT)
      r.f32(tp, "prob", &o.constant_prob);
    else if (o.term_prob == PT_TERM_N_ITERS)
      r.u32(tp, "iters", &o.n_iters);
    else if (o.term_prob == PT_TERM_MULTIPLIER_FUNC) {
      r.f32(tp, "exp", &o.multiplier_func_exp);
      r.f32(tp, "min_prob", &o.multiplier_func_min_prob);
    } else
      r.f32(tp, "min_prob", &o.normalize_min_prob);
  }
  const YNode *rv = r.tagged(st, "rng", kRngNames, 2, &o.rng);
  if (rv)
    r.u32(rv, "max_sample_size", &o.max_sample_size);
  r.boolean(st->get("rendering_equation_settings"), "back_cull_emission",
            &o.back_cull_emission);
  if (!r.ok)
    return false;
  *s = o;
  return true;
}

std::string settings_to_yaml(const pt_settings &s) {
  std::ostringstream o;
  auto wd = [&](const char *ind, const char *name, const pt_work_division_settings &w) {
    o << ind << name << ":\n";
    o << ind << "  block_size: " << w.block_size << "\n";
    o << ind << "  target_x_block_size: " << w.target_x_block_size << "\n";
    o << ind << "  force_target_samples: " << (w.force_target_samples ? "true" : "false") << "\n";
    o << ind << "  forced_target_samples_per_thread: " << w.forced_target_samples_per_thread << "\n";
    o << ind << "  base_num_threads: " << w.base_num_threads << "\n";
    o << ind << "  samples_per_thread_scaling_power: " << w.samples_per_thread_scaling_power << "\n";
    o << ind << "  max_samples_per_thread: " << w.max_samples_per_thread << "\n";
  };
  auto accel = [&](const char *ind) {
    o << ind << "accel:\n" << ind << "  type: " << kAccel[s.accel % 4] << "\n";
    if (s.accel == PT_ACCEL_SBVH) {
      o << ind << "  value:\n";
      o << ind << "    overlap_threshold: " << s.overlap_threshold << "\n";
      o << ind << "    use_spatial_splits: " << (s.use_spatial_splits ? "true" : "false") << "\n";
      o << ind << "    bvh_settings:\n";
      o << ind << "      traversal_per_intersect_cost: " << s.traversal_per_intersect_cost << "\n";
      o << ind << "      target_objects: " << s.target_objects << "\n";
      o << ind << "      print_stats: " << (s.print_stats ? "true" : "false") << "\n";
    } else {
      o << ind << "  value: {}\n";
    }
  };
  o << "settings:\n  kernel_approach:\n    type: " << kApproach[s.kernel_approach % 2]
    << "\n    value:\n      computation_settings:\n";
  if (s.kernel_approach == PT_MEGA_KERNEL) {
    o << "        max_blocks_per_launch_gpu: " << s.max_blocks_per_launch_gpu << "\n";
    o << "        max_blocks_per_launch_cpu: " << s.max_blocks_per_launch_cpu << "\n";
    wd("        ", "render_work_division", s.render_work_division);
    wd("        ", "reduce_work_division", s.reduce_work_division);
    o << "      individually_intersectable_settings:\n";
    accel("        ");
  } else {
    wd("        ", "init_samples_division", s.init_samples_division);
    o << "        max_num_samples_per_launch: " << s.max_num_samples_per_launch << "\n";
    o << "        min_num_init_blocks: " << s.min_num_init_blocks << "\n";
    o << "      accel:\n        type: IndividualToBulk\n        value:\n";
    o << "          to_bulk_settings:\n            max_size: " << s.to_bulk_max_size << "\n";
    o << "          individual_settings:\n";
    accel("            ");
  }
  o << "  light_sampler:\n    type: " << kLight[s.light_sampler % 2] << "\n";
  if (s.light_sampler == PT_LIGHT_RANDOM_TRIANGLE)
    o << "    value:\n      binary_search_threshold: " << s.binary_search_threshold << "\n";
  else
    o << "    value: {}\n";
  o << "  dir_sampler:\n    type: " << kDir[s.dir_sampler % 2] << "\n    value: {}\n";
  o << "  term_prob:\n    type: " << kTerm[s.term_prob % 4] << "\n    value:\n";
  switch (s.term_prob) {
  case PT_TERM_CONSTANT:
    o << "      prob: " << s.constant_prob << "\n";
    break;
  case PT_TERM_N_ITERS:
    o << "      iters: " << s.n_iters << "\n";
    break;
  case PT_TERM_MULTIPLIER_FUNC:
    o << "      exp: " << s.multiplier_func_exp << "\n      min_prob: "
      << s.multiplier_func_min_prob << "\n";
    break;
  default:
    o << "      min_prob: " << s.normalize_min_prob << "\n";
  }
  o << "  rng:\n    type: " << kRngNames[s.rng % 2]
    << "\n    value:\n      sequence_settings: {}\n      max_sample_size: "
    << s.max_sample_size << "\n";
  o << "  rendering_equation_settings:\n    back_cull_emission: "
    << (s.back_cull_emission ? "true" : "false") << "\n";
  return o.str();
}

} // namespace host
} // namespace pt
