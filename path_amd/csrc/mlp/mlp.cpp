// Radiance MLP host side: parameters, forward / backward schedules, Adam,
// resident training set, C ABI.
#include "mlp.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace pt;
using namespace pt::host;
using namespace pt::mlp;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error(PT_ERR_HIP, std::string(#expr) + ": " +                 \
                                       hipGetErrorString(e_));                 \
  } while (0)

pt_mlp::~pt_mlp() {
  if (ev0)
    (void)hipEventDestroy(ev0);
  if (ev1)
    (void)hipEventDestroy(ev1);
  if (stream)
    (void)hipStreamDestroy(stream);
}

// He-uniform initialisation with a seeded SplitMix64; padded input columns of
// layer 0 and padded output rows of the last layer are zero.
static std::vector<float> init_weights(uint64_t seed) {
  std::vector<float> w(param_count(), 0.f);
  uint64_t s = seed * 0x2545f4914f6cdd1dull + 0x51ed27ull;
  auto next = [&]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1p-53;
  };
  for (uint32_t l = 0; l < kLayers; ++l) {
    const uint32_t in = layer_in(l), out = layer_out(l);
    const uint32_t in_valid = l == 0 ? kEncoded : in;
    const uint32_t out_valid = l + 1 == kLayers ? kOut : out;
    const double bound = std::sqrt(6.0 / in_valid) * (l + 1 == kLayers ? 0.1 : 1.0);
    float *wl = w.data() + layer_offset(l);
    for (uint32_t o = 0; o < out_valid; ++o)
      for (uint32_t i = 0; i < in_valid; ++i)
        wl[size_t(o) * in + i] = float((2.0 * next() - 1.0) * bound);
  }
  return w;
}

bool pt_mlp::init(int dev, uint64_t seed, std::string *err) {
  device = dev;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipEventCreate(&ev0);
  if (e == hipSuccess)
    e = hipEventCreate(&ev1);
  const uint64_t P = param_count();
  std::vector<float> w = init_weights(seed);
  if (e == hipSuccess)
    e = w32.upload(w.data(), P, stream);
  if (e == hipSuccess)
    e = wbf.ensure(P);
  if (e == hipSuccess)
    e = launch_to_bf16(w32.ptr, wbf.ptr, P, stream);
  if (e == hipSuccess)
    e = wbf_t.ensure(P);
  if (e == hipSuccess)
    e = launch_transpose_weights(wbf.ptr, wbf_t.ptr, stream);
  if (e == hipSuccess)
    e = adam_m.ensure(P);
  if (e == hipSuccess)
    e = adam_v.ensure(P);
  if (e == hipSuccess)
    e = grad.ensure(P);
  if (e == hipSuccess)
    e = hipMemsetAsync(adam_m.ptr, 0, P * 4, stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(adam_v.ptr, 0, P * 4, stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    *err = std::string("MLP init: ") + hipGetErrorString(e);
    return false;
  }
  return true;
}

int pt_mlp::ensure_rows(uint32_t rows_pad, bool training) {
  HIP_TRY(feat.ensure(size_t(rows_pad) * kFeatures));
  HIP_TRY(targ.ensure(size_t(rows_pad) * kOut));
  HIP_TRY(pred.ensure(size_t(rows_pad) * kOutPad));
  HIP_TRY(x0.ensure(size_t(rows_pad) * kInPad));
  // inference ping-pongs act[0] / act[1]; training keeps every layer
  const uint32_t n_act = training ? kHidden : 2;
  for (uint32_t l = 0; l < n_act; ++l)
    HIP_TRY(act[l].ensure(size_t(rows_pad) * kWidth));
  if (training) {
    HIP_TRY(d_pred.ensure(size_t(rows_pad) * kOutPad));
    HIP_TRY(g0.ensure(size_t(rows_pad) * kWidth));
    HIP_TRY(g1.ensure(size_t(rows_pad) * kWidth));
    HIP_TRY(loss_part.ensure((rows_pad + 255) / 256));
    HIP_TRY(x0_t.ensure(size_t(rows_pad) * kInPad));
    for (uint32_t l = 0; l < kHidden; ++l)
      HIP_TRY(act_t[l].ensure(size_t(rows_pad) * kWidth));
    HIP_TRY(d_pred_t.ensure(size_t(rows_pad) * kOutPad));
    HIP_TRY(g0_t.ensure(size_t(rows_pad) * kWidth));
    HIP_TRY(g1_t.ensure(size_t(rows_pad) * kWidth));
  }
  return PT_OK;
}

static GemmArgs gemm_args(const bf16 *A, uint32_t lda, const bf16 *B, uint32_t ldb,
                          uint32_t M, uint32_t N, uint32_t K, void *C, uint32_t ldc) {
  GemmArgs g{};
  g.A = A;
  g.B = B;
  g.lda = lda;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  g.C = C;
  g.ldc = ldc;
  g.k_per_split = (K + 31) / 32 * 32;
  return g;
}

int pt_mlp::forward_device(const float *d_feat, uint32_t rows_pad, bool training,
                           hipStream_t on) {
  hipStream_t stream = on ? on : this->stream;
  if (!training && fused_forward) {
    HIP_TRY(launch_mlp_forward(d_feat, rows_pad, wbf.ptr, pred.ptr, stream));
    return PT_OK;
  }
  HIP_TRY(launch_encode(d_feat, rows_pad, x0.ptr, stream));
  if (training)
    HIP_TRY(launch_transpose(x0.ptr, rows_pad, kInPad, kInPad, x0_t.ptr, rows_pad, stream));
  const bf16 *in = x0.ptr;
  for (uint32_t l = 0; l < kLayers; ++l) {
    const bf16 *w = wbf.ptr + layer_offset(l);
    if (l + 1 < kLayers) {
      bf16 *out = training ? act[l].ptr : act[l & 1].ptr;
      GemmArgs g = gemm_args(in, layer_in(l), w, layer_in(l), rows_pad, kWidth, layer_in(l),
                             out, kWidth);
      if (training) {
        g.CT = act_t[l].ptr;
        g.ldct = rows_pad;
      }
      HIP_TRY(launch_gemm(g, false, false, kEpiReluBf16, 1, stream));
      in = out;
    } else {
      GemmArgs g = gemm_args(in, kWidth, w, kWidth, rows_pad, kOutPad, kWidth, pred.ptr,
                             kOutPad);
      HIP_TRY(launch_gemm(g, false, false, kEpiF32, 1, stream));
    }
  }
  return PT_OK;
}

int pt_mlp::backward(uint32_t n, uint32_t rows_pad, float *loss, float *loss_dev) {
  uint32_t n_part = 0;
  HIP_TRY(launch_loss(pred.ptr, targ.ptr, n, rows_pad, d_pred.ptr, d_pred_t.ptr, rows_pad,
                      loss_part.ptr, &n_part, stream));
  if (loss_dev && n_part)
    HIP_TRY(launch_loss_sum(loss_part.ptr, n_part, loss_dev, stream));
  // split-K over the batch for the weight gradients (512 rows per split, at
  // most 128 splits: 512 workgroups for a 256 x 256 gradient)
  const uint32_t splits = std::max(1u, std::min(128u, rows_pad / 512u));
  const uint32_t kps = (rows_pad + splits - 1) / splits;
  const uint32_t kps32 = (kps + 63) / 64 * 64;
  HIP_TRY(split_part.ensure(size_t(splits) * kWidth * kWidth));
  // G: gradient w.r.t. the layer's output [rows x out]; G_t the same
  // transposed [out x rows]
  const bf16 *G = d_pred.ptr, *G_t = d_pred_t.ptr;
  for (int l = int(kLayers) - 1; l >= 0; --l) {
    const uint32_t in = layer_in(l), out = layer_out(l);
    const bf16 *a_in_t = l == 0 ? x0_t.ptr : act_t[l - 1].ptr;
    // dW_l[out x in] = sum over rows of G_t[out][row] a_in_t[in][row]
    GemmArgs gw = gemm_args(G_t, rows_pad, a_in_t, rows_pad, out, in, rows_pad, split_part.ptr,
                            in);
    gw.k_per_split = kps32;
    const uint32_t used = (rows_pad + kps32 - 1) / kps32;
    HIP_TRY(launch_gemm(gw, false, false, kEpiSplitF32, used, stream));
    HIP_TRY(launch_sum_splits(split_part.ptr, used, uint64_t(out) * in,
                              grad.ptr + layer_offset(l), stream));
    if (l == 0)
      break;
    // dA[rows x in] = (G[rows x out] . W_l[out x in]) masked by the layer
    // input > 0 (read from its transposed copy, 4 rows per load); W_l^T
    // ([in][out]) is the B operand; also stored transposed
    const bool to0 = G != g0.ptr;
    bf16 *gn = to0 ? g0.ptr : g1.ptr;
    bf16 *gn_t = to0 ? g0_t.ptr : g1_t.ptr;
    GemmArgs gd = gemm_args(G, out, wbf_t.ptr + layer_offset(l), out, rows_pad, in, out, gn, in);
    gd.mask_t = act_t[l - 1].ptr;
    gd.ldmt = rows_pad;
    gd.CT = gn_t;
    gd.ldct = rows_pad;
    HIP_TRY(launch_gemm(gd, false, false, kEpiMaskBf16, 1, stream));
    G = gn;
    G_t = gn_t;
  }
  if (loss) {
    std::vector<float> parts(n_part);
    HIP_TRY(hipMemcpyAsync(parts.data(), loss_part.ptr, n_part * sizeof(float),
                           hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    double s = 0.0;
    for (float p : parts)
      s += p;
    *loss = float(s);
  }
  return PT_OK;
}

int pt_mlp::adam(float lr) {
  ++step;
  HIP_TRY(launch_adam(w32.ptr, grad.ptr, adam_m.ptr, adam_v.ptr, wbf.ptr, param_count(), lr,
                      0.9f, 0.99f, 1e-8f, step, stream));
  HIP_TRY(launch_transpose_weights(wbf.ptr, wbf_t.ptr, stream));
  return PT_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int pt_mlp_create(int device, uint64_t seed, pt_mlp **out) {
  if (!out)
    return set_error(PT_ERR_INVALID, "null output pointer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return set_error(PT_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= count)
    return set_error(PT_ERR_INVALID, "device index out of range");
  auto m = new (std::nothrow) pt_mlp();
  if (!m)
    return set_error(PT_ERR_INVALID, "out of host memory");
  std::string err;
  if (!m->init(device, seed, &err)) {
    delete m;
    return set_error(PT_ERR_HIP, err);
  }
  *out = m;
  return PT_OK;
}

void pt_mlp_destroy(pt_mlp *m) { delete m; }

int pt_mlp_param_count(uint64_t *count) {
  if (!count)
    return set_error(PT_ERR_INVALID, "null argument");
  *count = param_count();
  return PT_OK;
}

int pt_mlp_get_weights(pt_mlp *m, float *w) {
  if (!m || !w)
    return set_error(PT_ERR_INVALID, "null argument");
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipMemcpyAsync(w, m->w32.ptr, param_count() * 4, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return PT_OK;
}

int pt_mlp_set_weights(pt_mlp *m, const float *w) {
  if (!m || !w)
    return set_error(PT_ERR_INVALID, "null argument");
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(m->w32.upload(w, param_count(), m->stream));
  HIP_TRY(launch_to_bf16(m->w32.ptr, m->wbf.ptr, param_count(), m->stream));
  HIP_TRY(launch_transpose_weights(m->wbf.ptr, m->wbf_t.ptr, m->stream));
  HIP_TRY(hipMemsetAsync(m->adam_m.ptr, 0, param_count() * 4, m->stream));
  HIP_TRY(hipMemsetAsync(m->adam_v.ptr, 0, param_count() * 4, m->stream));
  m->step = 0;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return PT_OK;
}

int pt_mlp_set_forward_kernel(pt_mlp *m, int fused) {
  if (!m)
    return set_error(PT_ERR_INVALID, "null argument");
  m->fused_forward = fused != 0;
  return PT_OK;
}

int pt_mlp_forward(pt_mlp *m, const float *features, uint32_t n, float *out) {
  if (!m || (n && (!features || !out)))
    return set_error(PT_ERR_INVALID, "null argument");
  if (n == 0)
    return PT_OK;
  HIP_TRY(hipSetDevice(m->device));
  const uint32_t rp = mlp_pad_rows(n);
  int rc = m->ensure_rows(rp, false);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(hipMemsetAsync(m->feat.ptr, 0, size_t(rp) * kFeatures * 4, m->stream));
  HIP_TRY(hipMemcpyAsync(m->feat.ptr, features, size_t(n) * kFeatures * 4,
                         hipMemcpyHostToDevice, m->stream));
  rc = m->forward_device(m->feat.ptr, rp, false);
  if (rc != PT_OK)
    return rc;
  // compact to n x 3 on the device (reuses the front of the target buffer)
  HIP_TRY(m->targ.ensure(size_t(rp) * kOut));
  HIP_TRY(launch_compact_pred(m->pred.ptr, n, m->targ.ptr, m->stream));
  HIP_TRY(hipMemcpyAsync(out, m->targ.ptr, size_t(n) * kOut * 4, hipMemcpyDeviceToHost,
                         m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return PT_OK;
}

static int load_batch(pt_mlp *m, const float *features, const float *targets, uint32_t n,
                      uint32_t *rp_out) {
  const uint32_t rp = mlp_pad_rows(n);
  int rc = m->ensure_rows(rp, true);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(hipMemsetAsync(m->feat.ptr, 0, size_t(rp) * kFeatures * 4, m->stream));
  HIP_TRY(hipMemcpyAsync(m->feat.ptr, features, size_t(n) * kFeatures * 4,
                         hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->targ.ptr, targets, size_t(n) * kOut * 4, hipMemcpyHostToDevice,
                         m->stream));
  *rp_out = rp;
  return PT_OK;
}

int pt_mlp_gradient(pt_mlp *m, const float *features, const float *targets, uint32_t n,
                    float *grad, float *loss) {
  if (!m || !features || !targets || !grad || n == 0)
    return set_error(PT_ERR_INVALID, "null argument or empty batch");
  HIP_TRY(hipSetDevice(m->device));
  uint32_t rp;
  int rc = load_batch(m, features, targets, n, &rp);
  if (rc == PT_OK)
    rc = m->forward_device(m->feat.ptr, rp, true);
  if (rc == PT_OK)
    rc = m->backward(n, rp, loss);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(hipMemcpyAsync(grad, m->grad.ptr, param_count() * 4, hipMemcpyDeviceToHost,
                         m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return PT_OK;
}

int pt_mlp_train_step(pt_mlp *m, const float *features, const float *targets, uint32_t n,
                      float lr, float *loss) {
  if (!m || !features || !targets || n == 0)
    return set_error(PT_ERR_INVALID, "null argument or empty batch");
  HIP_TRY(hipSetDevice(m->device));
  uint32_t rp;
  int rc = load_batch(m, features, targets, n, &rp);
  if (rc == PT_OK)
    rc = m->forward_device(m->feat.ptr, rp, true);
  if (rc == PT_OK)
    rc = m->backward(n, rp, loss);
  if (rc == PT_OK)
    rc = m->adam(lr);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return PT_OK;
}

// Times `iters` inference passes over `rows` device-resident rows (features
// all 0.25): the average duration of one hidden 256x256 layer GEMM (layer 1)
// and of the whole forward pass, both with HIP events on the MLP's stream.
int pt_mlp_time_forward(pt_mlp *m, uint32_t rows, uint32_t iters, float *gemm_ms,
                        float *forward_ms) {
  if (!m || rows == 0 || iters == 0)
    return set_error(PT_ERR_INVALID, "null argument or empty workload");
  HIP_TRY(hipSetDevice(m->device));
  const uint32_t rp = mlp_pad_rows(rows);
  int rc = m->ensure_rows(rp, false);
  if (rc != PT_OK)
    return rc;
  std::vector<float> f(size_t(rp) * kFeatures, 0.25f);
  HIP_TRY(hipMemcpyAsync(m->feat.ptr, f.data(), f.size() * 4, hipMemcpyHostToDevice,
                         m->stream));
  rc = m->forward_device(m->feat.ptr, rp, false);
  if (rc != PT_OK)
    return rc;
  GemmArgs g = gemm_args(m->act[0].ptr, kWidth, m->wbf.ptr + layer_offset(1), kWidth, rp,
                         kWidth, kWidth, m->act[1].ptr, kWidth);
  HIP_TRY(hipEventRecord(m->ev0, m->stream));
  for (uint32_t i = 0; i < iters; ++i)
    HIP_TRY(launch_gemm(g, false, false, kEpiReluBf16, 1, m->stream));
  HIP_TRY(hipEventRecord(m->ev1, m->stream));
  HIP_TRY(hipEventSynchronize(m->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  if (gemm_ms)
    *gemm_ms = ms / float(iters);
  HIP_TRY(hipEventRecord(m->ev0, m->stream));
  for (uint32_t i = 0; i < iters && rc == PT_OK; ++i)
    rc = m->forward_device(m->feat.ptr, rp, false);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(hipEventRecord(m->ev1, m->stream));
  HIP_TRY(hipEventSynchronize(m->ev1));
  HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  if (forward_ms)
    *forward_ms = ms / float(iters);
  return PT_OK;
}

int pt_mlp_set_dataset(pt_mlp *m, const float *features, const float *targets, uint64_t rows) {
  if (!m || (rows && (!features || !targets)))
    return set_error(PT_ERR_INVALID, "null argument");
  if (rows >= (1ull << 32))
    return set_error(PT_ERR_INVALID, "at most 2^32 - 1 training rows");
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(m->ds_feat.upload(features, rows * kFeatures, m->stream));
  HIP_TRY(m->ds_targ.upload(targets, rows * kOut, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->ds_rows = rows;
  return PT_OK;
}

int pt_mlp_train(pt_mlp *m, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                 float *losses, double *seconds) {
  if (!m || batch == 0)
    return set_error(PT_ERR_INVALID, "null argument or empty batch");
  if (m->ds_rows == 0)
    return set_error(PT_ERR_INVALID, "no training set (pt_mlp_set_dataset)");
  HIP_TRY(hipSetDevice(m->device));
  const uint32_t rp = mlp_pad_rows(batch);
  int rc = m->ensure_rows(rp, true);
  if (rc != PT_OK)
    return rc;
  HIP_TRY(m->idx.ensure(rp));
  if (losses)
    HIP_TRY(m->step_losses.ensure(std::max<uint32_t>(steps, 1)));
  HIP_TRY(hipMemsetAsync(m->feat.ptr, 0, size_t(rp) * kFeatures * 4, m->stream));
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t it = 0; it < steps; ++it) {
    HIP_TRY(launch_random_rows(batch, m->ds_rows, seed + it * 0x632be59bd9b4e019ull,
                               m->idx.ptr, m->stream));
    HIP_TRY(launch_gather(m->ds_feat.ptr, m->ds_targ.ptr, m->idx.ptr, batch, m->feat.ptr,
                          m->targ.ptr, m->stream));
    rc = m->forward_device(m->feat.ptr, rp, true);
    if (rc == PT_OK)
      rc = m->backward(batch, rp, nullptr, losses ? m->step_losses.ptr + it : nullptr);
    if (rc == PT_OK)
      rc = m->adam(lr);
    if (rc != PT_OK)
      return rc;
  }
  // the per-step losses stay on the device until the end: no step waits
  // for the host
  if (losses && steps)
    HIP_TRY(hipMemcpyAsync(losses, m->step_losses.ptr, size_t(steps) * sizeof(float),
                           hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PT_OK;
}

} // extern "C"
