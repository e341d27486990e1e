// Radiance MLP on MFMA: BF16 operands, FP32 accumulation.
//
// Network: 9 features (position, shading normal, albedo) -> 12-frequency
// encoding (9 raw + 9 x 12 x {sin, cos} = 225, zero-padded to 256) -> 6
// hidden layers of width 256 with ReLU -> 3 outputs (padded to 16). Weights
// are [out][in] row-major; no biases.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>

namespace pt {
namespace mlp {

using bf16 = __hip_bfloat16;

constexpr uint32_t kFeatures = 9;
constexpr uint32_t kFrequencies = 12;
constexpr uint32_t kEncoded = kFeatures * (1 + 2 * kFrequencies); // 225
constexpr uint32_t kInPad = 256;
constexpr uint32_t kWidth = 256;
constexpr uint32_t kHidden = 6;
constexpr uint32_t kOut = 3;
constexpr uint32_t kOutPad = 16;
constexpr uint32_t kLayers = kHidden + 1; // 7 weight matrices
constexpr uint32_t kRowAlign = 128;       // batch rows are a multiple of this

inline uint32_t layer_in(uint32_t l) { return l == 0 ? kInPad : kWidth; }
inline uint32_t layer_out(uint32_t l) { return l + 1 == kLayers ? kOutPad : kWidth; }
inline uint64_t layer_offset(uint32_t l) {
  uint64_t o = 0;
  for (uint32_t i = 0; i < l; ++i)
    o += uint64_t(layer_in(i)) * layer_out(i);
  return o;
}
inline uint64_t param_count() { return layer_offset(kLayers); }

enum Epilogue : int {
  kEpiReluBf16 = 0, // C = bf16(max(acc, 0))
  kEpiF32 = 1,      // C = acc (fp32)
  kEpiMaskBf16 = 2, // C = bf16(acc * (layer input > 0)), the input read transposed
  kEpiSplitF32 = 3, // C[z] = acc (fp32 partial of K split z)
};

// C[M x N] = sum_k A(m, k) B(k, n) over k in this split's range.
// A(m, k) = a_trans ? A[k * lda + m] : A[m * lda + k]
// B(k, n) = b_trans ? B[k * ldb + n] : B[n * ldb + k]
// Requirements: M, N, K, lda, ldb multiples of 8; k_per_split a multiple of 32.
// With the bf16 epilogues, CT (when set) also receives the result
// transposed: CT[n * ldct + m] (M a multiple of 8, ldct of 4), the layout in
// which the weight-gradient GEMM reads it with the batch as its K.
// kEpiMaskBf16 reads its mask transposed: element (m, n) at
// mask_t[n * ldmt + m] (ldmt a multiple of 4), 4 rows per load.
struct GemmArgs {
  const bf16 *A;
  const bf16 *B;
  uint32_t lda, ldb;
  uint32_t M, N, K;
  void *C;
  uint32_t ldc;
  uint32_t k_per_split;
  void *CT;
  uint32_t ldct;
  const bf16 *mask_t;
  uint32_t ldmt;
};

hipError_t launch_gemm(const GemmArgs &g, bool a_trans, bool b_trans, int epilogue,
                       uint32_t splits, hipStream_t s);
// features [n x 9] fp32 -> encoded [n x 256] bf16
hipError_t launch_encode(const float *features, uint32_t n, bf16 *out, hipStream_t s);
// Whole inference forward in one kernel (encoding, hidden layers, output):
// features [rows_pad x 9] -> pred [rows_pad x 16] fp32, rows_pad a multiple
// of 128; equal bit for bit to encode + the layer GEMMs.
hipError_t launch_mlp_forward(const float *features, uint32_t rows_pad, const bf16 *w,
                              float *pred, hipStream_t s);
// Relative L2 loss on the first 3 outputs: l = (p - t)^2 / (p^2 + 0.01) with p
// held constant in the denominator, mean over n * 3. Writes d(loss)/d(pred)
// as bf16 [n x 16] (padding zero) and per-block loss partials (fp32).
// Rows n .. rows_pad - 1 of d_pred are zeroed; d_pred_t (optional) receives
// d_pred transposed: [16 x ld_t].
hipError_t launch_loss(const float *pred, const float *target, uint32_t n, uint32_t rows_pad,
                       bf16 *d_pred, bf16 *d_pred_t, uint32_t ld_t, float *loss_partials,
                       uint32_t *n_partials, hipStream_t s);
// *out = sum of the n loss partials in order (FP64), on the device.
hipError_t launch_loss_sum(const float *partials, uint32_t n, float *out, hipStream_t s);
// dst[c * ld_dst + r] = src[r * ld_src + c] for r < rows, c < cols (bf16).
hipError_t launch_transpose(const bf16 *src, uint32_t rows, uint32_t cols, uint32_t ld_src,
                            bf16 *dst, uint32_t ld_dst, hipStream_t s);
// Per layer, the bf16 weights transposed ([in][out]): the B operand of the
// data-gradient GEMM. One launch for all layers.
hipError_t launch_transpose_weights(const bf16 *w, bf16 *w_t, hipStream_t s);
// out[i] = sum over z < splits of partials[z * count + i] (fixed order)
hipError_t launch_sum_splits(const float *partials, uint32_t splits, uint64_t count,
                             float *out, hipStream_t s);
// Adam (bias-corrected with step t >= 1) on fp32 master weights; also writes
// the bf16 copy used by the GEMMs.
hipError_t launch_adam(float *w, const float *g, float *m, float *v, bf16 *w_bf16,
                       uint64_t count, float lr, float beta1, float beta2, float eps,
                       uint32_t t, hipStream_t s);
hipError_t launch_to_bf16(const float *in, bf16 *out, uint64_t count, hipStream_t s);
// out[i][c] = pred[i][c] for c < 3 (rows of 16 -> rows of 3)
hipError_t launch_compact_pred(const float *pred, uint32_t n, float *out, hipStream_t s);
// Gather rows idx[i] of a dataset (features 9 + targets 3 per row).
hipError_t launch_gather(const float *feat, const float *targ, const uint32_t *idx,
                         uint32_t n, float *feat_out, float *targ_out, hipStream_t s);
// idx[i] = hash(seed, i) % n_rows
hipError_t launch_random_rows(uint32_t n, uint64_t n_rows, uint64_t seed, uint32_t *idx,
                              hipStream_t s);

} // namespace mlp
} // namespace pt
