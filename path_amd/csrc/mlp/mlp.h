// Radiance MLP state behind the C ABI.
#pragma once

#include "../host/renderer.h"
#include "mlp_kernels.h"

struct pt_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t step = 0; // Adam steps taken
  bool fused_forward = true; // inference through k_mlp_forward

  pt::host::DeviceBuf<float> w32, adam_m, adam_v, grad;
  // bf16 weights, and per layer transposed ([in][out]) for the data gradient
  pt::host::DeviceBuf<pt::mlp::bf16> wbf, wbf_t;

  // per-batch buffers (grow-only, rows padded to kRowAlign)
  pt::host::DeviceBuf<float> feat, targ, pred, loss_part, split_part;
  pt::host::DeviceBuf<pt::mlp::bf16> x0, act[pt::mlp::kHidden], d_pred, g0, g1;
  // training: the same tensors transposed ([column][row]), the operands of
  // the weight-gradient GEMMs, whose K is the batch
  pt::host::DeviceBuf<pt::mlp::bf16> x0_t, act_t[pt::mlp::kHidden], d_pred_t, g0_t, g1_t;
  pt::host::DeviceBuf<uint32_t> idx;
  pt::host::DeviceBuf<float> step_losses; // pt_mlp_train: one mean loss per step

  // resident training set
  pt::host::DeviceBuf<float> ds_feat, ds_targ;
  uint64_t ds_rows = 0;

  ~pt_mlp();
  bool init(int dev, uint64_t seed, std::string *err);
  int ensure_rows(uint32_t rows_pad, bool training);
  // Forward on device features (rows_pad rows, padded rows arbitrary):
  // pred[rows_pad x 16] fp32. training = keep every layer's activations
  // (layer-by-layer GEMMs); otherwise the fused kernel unless disabled.
  int forward_device(const float *d_feat, uint32_t rows_pad, bool training,
                     hipStream_t on = nullptr);
  // Loss + backward + gradient into `grad` for the batch already in
  // feat/targ (n valid rows of rows_pad). The mean loss goes to *loss (host,
  // synchronising) and/or *loss_dev (device, no synchronisation).
  int backward(uint32_t n, uint32_t rows_pad, float *loss, float *loss_dev = nullptr);
  int adam(float lr);
};

namespace pt {
namespace host {
inline uint32_t mlp_pad_rows(uint64_t n) {
  return uint32_t((n + mlp::kRowAlign - 1) / mlp::kRowAlign * mlp::kRowAlign);
}
} // namespace host
} // namespace pt
