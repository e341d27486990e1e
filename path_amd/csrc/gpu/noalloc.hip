// Mega-kernel render into caller device memory that allocates nothing, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid it).
//
// The render's scratch memory - the ray counter and, when a pixel's samples
// span several sample blocks, the per-block partial sums - lives in one
// file-scope device array per GPU instead of renderer-owned allocations. The
// scene and the Sobol table must already be resident on the renderer (an
// earlier pt_render of the same scene with the same samples_per and
// max_sample_size leaves them there); nothing is uploaded or allocated.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents: the counter is cleared by a kernel at the start of each call, and
// calls that overlap on one GPU (two renderers on two host threads) race.
#include "../host/renderer.h"
#include "../host/settings.h"
#include "kernels.h"

#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

namespace {

// 16 MiB: 256 bytes for the ray counter, then up to (16 MiB - 256) / 12
// pixel x sample-block partial sums (FloatRGB).
constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr size_t kCounterBytes = 256;
constexpr size_t kMaxPartials = (kScratchBytes - kCounterBytes) / (3 * sizeof(float));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU (each GPU holds its own copy of the
// array at its own address); filled once per device at renderer set-up.
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

__global__ void k_clear_counter(unsigned long long *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *counter = 0ull;
}

} // namespace

namespace pt {
namespace host {

// Called by pt_renderer::init with `device` current.
int noalloc_init(int device, std::string *err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = "device index out of range for the no-allocation scratch table";
    return PT_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device])
    return PT_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return PT_ERR_HIP;
  }
  g_scratch_ptr[device] = ptr;
  return PT_OK;
}

} // namespace host
} // namespace pt

using namespace pt;
using namespace pt::host;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error(PT_ERR_HIP, std::string(#expr) + ": " +                 \
                                       hipGetErrorString(e_));                 \
  } while (0)

extern "C" int pt_render_noalloc(pt_renderer *r, const pt_sample_spec *spec,
                                 float *device_rgb, const pt_scene *scene,
                                 uint32_t samples_per, const pt_settings *settings,
                                 double *seconds) {
  try {
    if (!r || !spec || !device_rgb || !scene || !settings)
      return set_error(PT_ERR_INVALID, "null argument");
    if (spec->type != PT_SAMPLE_SQUARE_IMAGE)
      return set_error(PT_ERR_INVALID, "pt_render_noalloc renders SquareImage sample specs");
    if (settings->kernel_approach != PT_MEGA_KERNEL)
      return set_error(PT_ERR_INVALID, "pt_render_noalloc runs the MegaKernel integrator");
    if (samples_per == 0 || spec->x_dim == 0 || spec->y_dim == 0)
      return set_error(PT_ERR_INVALID, "empty image or samples_per = 0");
    void *base = r->device >= 0 && r->device < kMaxDevices ? g_scratch_ptr[r->device] : nullptr;
    if (!base)
      return set_error(PT_ERR_INVALID, "no-allocation scratch not set up on this device");
    // resident scene and sampler (set up by an earlier pt_render)
    if (r->cached_scene != scene || r->cached_gen != scene->generation ||
        r->cached_mat_gen != scene->mat_generation)
      return set_error(PT_ERR_INVALID,
                       "scene not resident on the renderer: render it once with pt_render "
                       "first (pt_render_noalloc uploads nothing)");
    if (settings->rng == PT_RNG_SOBEL &&
        (r->sobol_dim_bound != settings->max_sample_size || r->sobol_count != samples_per))
      return set_error(PT_ERR_INVALID,
                       "Sobol table for this samples_per / max_sample_size not resident: "
                       "render once with pt_render first");
    WorkDivision wd;
    if (!WorkDivision::make(work_division_settings(settings->render_work_division),
                            samples_per, spec->x_dim, spec->y_dim, &wd) ||
        wd.block_size > 1024)
      return set_error(PT_ERR_INVALID, "invalid render work division settings");
    const size_t n_pixels = size_t(spec->x_dim) * spec->y_dim;
    const bool blocks = wd.num_sample_blocks != 1;
    if (blocks && n_pixels * wd.num_sample_blocks > kMaxPartials)
      return set_error(PT_ERR_INVALID,
                       "image x sample blocks (" + std::to_string(n_pixels * wd.num_sample_blocks) +
                           ") above the no-allocation limit of " + std::to_string(kMaxPartials) +
                           " partial sums");
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, device_rgb) != hipSuccess ||
        attr.type != hipMemoryTypeDevice || attr.device != r->device) {
      (void)hipGetLastError();
      return set_error(PT_ERR_INVALID, "device_rgb must be device memory of the renderer's GPU");
    }
    HIP_TRY(hipSetDevice(r->device));
    unsigned long long *counter = static_cast<unsigned long long *>(base);
    float *partials = reinterpret_cast<float *>(static_cast<unsigned char *>(base) + kCounterBytes);

    gpu::MegaParams p{};
    p.scene = r->scene_store.dev;
    p.settings = integrator_settings(*settings);
    p.wd = wd;
    p.rng.type = settings->rng;
    p.rng.samples_per = samples_per;
    if (settings->rng == PT_RNG_SOBEL) {
      p.rng.vals = r->sobol_vals.ptr;
      p.rng.dim_bound = settings->max_sample_size;
      p.rng.initial_dim_bound = settings->max_sample_size < 64 ? settings->max_sample_size : 64;
    }
    std::memcpy(p.film_to_world, spec->film_to_world, sizeof(p.film_to_world));
    p.out_rgb = blocks ? partials : device_rgb;
    p.divide_by_spp = blocks ? 0u : 1u;
    p.ray_counter = counter;

    r->stats = pt_render_stats{};
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(r->ev_start, r->stream));
    hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(64), 0, r->stream, counter);
    HIP_TRY(hipGetLastError());
    const uint32_t total = wd.total_num_blocks();
    const uint32_t chunk =
        settings->max_blocks_per_launch_gpu ? settings->max_blocks_per_launch_gpu : total;
    for (uint32_t start = 0; start < total; start += chunk) {
      const uint32_t end = start + chunk < total ? start + chunk : total;
      HIP_TRY(gpu::launch_megakernel(p, start, end, r->mega_split, r->stream));
      r->stats.mega_launches++;
    }
    if (blocks)
      HIP_TRY(gpu::launch_reduce_sample_blocks(partials, wd.num_sample_blocks,
                                               uint32_t(n_pixels), samples_per, device_rgb,
                                               r->stream));
    HIP_TRY(hipEventRecord(r->ev_stop, r->stream));
    // The ray count comes back only through pinned memory the renderer
    // already holds (an earlier pt_render with a host output allocated it);
    // a copy into pageable memory could make the runtime stage it.
    const bool count_back = r->out_stage.ptr && r->out_stage.bytes >= sizeof(unsigned long long);
    if (count_back)
      HIP_TRY(hipMemcpyAsync(r->out_stage.ptr, counter, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    unsigned long long rays = 0;
    if (count_back)
      std::memcpy(&rays, r->out_stage.ptr, sizeof(rays));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r->ev_start, r->ev_stop));
    r->stats.kernel_ms = ms;
    r->stats.rays = rays;
    r->stats.samples = uint64_t(samples_per) * n_pixels;
    if (seconds)
      *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PT_OK;
  } catch (const std::exception &e) {
    return set_error(PT_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}
