// `path` command line: the flags and behaviour of the reference's src/main.cpp
// (usage text at src/main.cpp:16-46), rendering on the GPU through the C ABI.
// Images are written as 8-bit RGB PNG.
#include "../../../include/path_amd.h"
#include "../core/camera.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

const char kUsage[] = R"(Path

    Usage:
      path <scene_file> [--width=<pixels>] [--height=<pixels>]
        [--samples=<count>] [--output=<file_name>] [--n-output-steps=<steps>]
        [--output-per-step] [--bench-budget=<time>] [--config=<file_name>]
        [-g | --gpu] [--bench] [--bench-ignore-accel] [--disable-progress]
        [--show-times] [--no-print-config]
      path (-h | --help)

    Options:
      -h --help                 Show this screen.
      --width=<pixels>          Width in pixels [default: 1024]
      --height=<pixels>         Height in pixels [default: 1024]
      --samples=<count>         Samples per pixel [default: 128]
      --output=<file_name>      File name [default: out.png]
      --output-per-step         Create output image per each step
      --n-output-steps=<steps>  Number of steps for output images [default: 8]
      --bench-budget=<time>     Approx time in seconds for bench [default: 5.0]
      --config=<file_name>      Config file name. If no file is specified,
                                default settings will be used.
      -g --gpu                  Use gpu (always: path_amd renders on the GPU)

      --bench                   Warm up and then run multiple times and report
                                statistics
      --bench-ignore-accel      Ignore the computation used to build the accel
      --disable-progress        Disable progress bar
      --show-times              Show timings
      --no-print-config         Don't print config
)";

// --- minimal PNG writer (stored deflate blocks) -----------------------------
uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xffffffffu) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t v = i;
      for (int k = 0; k < 8; ++k)
        v = (v & 1u) ? 0xedb88320u ^ (v >> 1) : v >> 1;
      table[i] = v;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i)
    c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
  return c;
}

void put_u32(std::vector<uint8_t> &v, uint32_t x) {
  v.push_back(uint8_t(x >> 24));
  v.push_back(uint8_t(x >> 16));
  v.push_back(uint8_t(x >> 8));
  v.push_back(uint8_t(x));
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
  put_u32(out, uint32_t(data.size()));
  std::vector<uint8_t> td(type, type + 4);
  td.insert(td.end(), data.begin(), data.end());
  out.insert(out.end(), td.begin(), td.end());
  put_u32(out, crc32(td.data(), td.size()) ^ 0xffffffffu);
}

// bgra: memory order B, G, R, 0 per pixel (the BGRA output of pt_render)
bool write_png(const std::string &path, const uint32_t *bgra, uint32_t w, uint32_t h) {
  std::vector<uint8_t> raw;
  raw.reserve(size_t(h) * (1 + 3 * size_t(w)));
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0); // filter: none
    for (uint32_t x = 0; x < w; ++x) {
      const uint32_t v = bgra[size_t(y) * w + x];
      raw.push_back(uint8_t(v >> 16));
      raw.push_back(uint8_t(v >> 8));
      raw.push_back(uint8_t(v));
    }
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
    const size_t n = std::min<size_t>(65535, raw.size() - off);
    const bool last = off + n >= raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(uint8_t(n));
    z.push_back(uint8_t(n >> 8));
    z.push_back(uint8_t(~n));
    z.push_back(uint8_t(~n >> 8));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
    if (last)
      break;
  }
  put_u32(z, (b << 16) | a);
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  put_u32(ihdr, w);
  put_u32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0}); // 8-bit RGB
  chunk(png, "IHDR", ihdr);
  chunk(png, "IDAT", z);
  chunk(png, "IEND", {});
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f)
    return false;
  const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
  return std::fclose(f) == 0 && ok;
}

uint32_t to_bgra(const float *rgb) { return pt::float_rgb_to_bgra32(rgb[0], rgb[1], rgb[2]); }

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(1);
}

void check(int rc, const char *what) {
  if (rc != PT_OK)
    die(std::string(what) + ": " + pt_last_error());
}

} // namespace

int main(int argc, char **argv) {
  std::map<std::string, std::string> opt = {
      {"--width", "1024"},         {"--height", "1024"},      {"--samples", "128"},
      {"--output", "out.png"},     {"--n-output-steps", "8"}, {"--bench-budget", "5.0"}};
  std::map<std::string, bool> flag;
  std::string scene_file, config_file;
  const char *valued[] = {"--width", "--height", "--samples", "--output",
                          "--n-output-steps", "--bench-budget", "--config"};
  const char *flags[] = {"--output-per-step", "--gpu", "--bench", "--bench-ignore-accel",
                         "--disable-progress", "--show-times", "--no-print-config"};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      std::fputs(kUsage, stdout);
      return 0;
    }
    if (a == "-g")
      a = "--gpu";
    bool matched = false;
    for (const char *v : valued) {
      const std::string k = v;
      if (a == k || a.rfind(k + "=", 0) == 0) {
        std::string val;
        if (a == k) {
          if (i + 1 >= argc)
            die("missing value for " + k);
          val = argv[++i];
        } else {
          val = a.substr(k.size() + 1);
        }
        (k == "--config" ? config_file : opt[k]) = val;
        matched = true;
      }
    }
    for (const char *f : flags)
      if (a == f) {
        flag[a] = true;
        matched = true;
      }
    if (matched)
      continue;
    if (!a.empty() && a[0] == '-')
      die("unknown option " + a + "\n" + kUsage);
    if (!scene_file.empty())
      die(std::string("more than one scene file\n") + kUsage);
    scene_file = a;
  }
  if (scene_file.empty())
    die(kUsage);
  const uint32_t width = uint32_t(std::stoul(opt["--width"]));
  const uint32_t height = uint32_t(std::stoul(opt["--height"]));
  const uint32_t samples = uint32_t(std::stoul(opt["--samples"]));
  const uint32_t n_steps = uint32_t(std::stoul(opt["--n-output-steps"]));
  const double budget = std::stod(opt["--bench-budget"]);

  pt_scene *scene = nullptr;
  float ftw[12];
  check(pt_scene_load(scene_file.c_str(), float(width) / float(height), 0, &scene, ftw),
        "load_scene");
  pt_settings settings;
  pt_settings_default(&settings);
  if (!config_file.empty())
    check(pt_settings_load_yaml(config_file.c_str(), &settings), "load_config");
  if (!flag["--no-print-config"]) {
    size_t need = 0;
    pt_settings_to_yaml(&settings, nullptr, 0, &need);
    std::string text(need + 1, '\0');
    check(pt_settings_to_yaml(&settings, text.data(), text.size(), &need), "print_config");
    std::fputs(text.c_str(), stdout);
  }
  pt_renderer *r = nullptr;
  check(pt_renderer_create(0, &r), "renderer");

  pt_sample_spec spec{};
  spec.type = PT_SAMPLE_SQUARE_IMAGE;
  spec.x_dim = width;
  spec.y_dim = height;
  std::memcpy(spec.film_to_world, ftw, sizeof(ftw));
  std::vector<uint32_t> pixels(size_t(width) * height);
  pt_output out{};
  out.type = PT_OUTPUT_BGRA;
  out.data = pixels.data();
  const bool progress = !flag["--disable-progress"];
  const bool show_times = flag["--show-times"];
  auto render = [&](bool show_progress) {
    double secs = 0.0;
    check(pt_render(r, &spec, &out, scene, samples, &settings, show_progress && progress,
                    show_times, &secs),
          "render");
    return secs;
  };

  if (flag["--bench"]) {
    using clock = std::chrono::steady_clock;
    render(false);
    const unsigned warmup_iters = 2;
    auto t0 = clock::now();
    for (unsigned i = 0; i < warmup_iters; ++i)
      render(false);
    const double time_per =
        std::chrono::duration<double>(clock::now() - t0).count() / warmup_iters;
    const unsigned iters = unsigned(std::ceil(std::max(0.0, budget / time_per)));
    if (iters < 3)
      std::fprintf(stderr, "Note that only %u iter(s) can be run with the current budget\n",
                   iters);
    double total = 0.0, total_sqr = 0.0;
    for (unsigned i = 0; i < iters; ++i) {
      auto t1 = clock::now();
      const double ignore_accel = render(false);
      const double t = flag["--bench-ignore-accel"]
                           ? ignore_accel
                           : std::chrono::duration<double>(clock::now() - t1).count();
      total += t;
      total_sqr += t * t;
    }
    const double mean = total / iters;
    const double var = total_sqr / iters - mean * mean;
    const double se = std::sqrt(std::max(var, 0.0)) / std::sqrt(double(iters));
    std::printf("mean: %g, 2 * SE: %g\n", mean, 2 * se);
  } else {
    render(true);
  }

  if (flag["--output-per-step"]) {
    std::vector<std::vector<float>> steps(n_steps, std::vector<float>(size_t(width) * height * 3));
    std::vector<float *> ptrs(n_steps);
    for (uint32_t i = 0; i < n_steps; ++i)
      ptrs[i] = steps[i].data();
    pt_output ps{};
    ps.type = PT_OUTPUT_PER_STEP;
    ps.per_step = ptrs.data();
    ps.n_steps = n_steps;
    pt_settings mk = settings;
    mk.kernel_approach = PT_MEGA_KERNEL; // per-step output runs on the mega-kernel
    double secs;
    check(pt_render(r, &spec, &ps, scene, samples, &mk, progress, show_times, &secs),
          "render per step");
    std::vector<uint32_t> img(size_t(width) * height);
    std::vector<float> sum(size_t(width) * height * 3, 0.f);
    for (uint32_t i = 0; i < n_steps; ++i) {
      for (size_t p = 0; p < img.size(); ++p) {
        img[p] = to_bgra(&steps[i][3 * p]);
        for (int c = 0; c < 3; ++c)
          sum[3 * p + c] += steps[i][3 * p + c];
      }
      if (!write_png("out_" + std::to_string(i) + ".png", img.data(), width, height))
        die("cannot write per-step image");
    }
    for (size_t p = 0; p < img.size(); ++p)
      img[p] = to_bgra(&sum[3 * p]);
    if (!write_png("summed_image.png", img.data(), width, height))
      die("cannot write summed_image.png");
  }

  if (!write_png(opt["--output"], pixels.data(), width, height))
    die("cannot write " + opt["--output"]);
  pt_renderer_destroy(r);
  pt_scene_destroy(scene);
  return 0;
}
