// Radiance MLP kernels: MFMA GEMM (BF16 in, FP32 accumulate), frequency
// encoding, relative-L2 loss, split-K reduction, Adam.
#include "mlp_kernels.h"

namespace pt {
namespace mlp {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float x) {
  return __builtin_bit_cast(uint16_t, __bf16(x));
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// ---------------------------------------------------------------------------
// GEMM: 128 x 128 block tile, K step 64, 4 waves in a 2 x 2 grid, each wave a
// 64 x 64 tile of 2 x 2 v_mfma_f32_32x32x16_bf16. Operands are staged in LDS
// as [row][k] (k contiguous, row stride 72 bf16 = 144 B); transposed global
// layouts are transposed on the way in. The next K step's global loads are
// issued into registers before the current step's MFMAs, so their latency
// overlaps the math.
// MFMA lane layout (32x32x16): lane l holds A[l % 32][8 (l / 32) .. +8],
// B[8 (l / 32) .. +8][l % 32]; accumulator r of lane l is
// C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// ---------------------------------------------------------------------------
constexpr uint32_t kBM = 128, kBN = 128, kBK = 64, kLdsStride = 72;
constexpr uint32_t kChunks = kBM * kBK / 8 / 256; // 16-byte chunks per thread per operand

// Global -> registers: chunk c of the tile (8 bf16 along k, or along rows when
// transposed), zero outside [0, rows) x [k0, k_end).
template <bool kT>
__device__ __forceinline__ void fetch_tile(const uint16_t *__restrict__ src, uint32_t ld,
                                           uint32_t r0, uint32_t rows, uint32_t k0,
                                           uint32_t k_end, uint4 (&v)[kChunks]) {
#pragma unroll
  for (uint32_t i = 0; i < kChunks; ++i) {
    const uint32_t c = threadIdx.x + 256u * i;
    v[i] = make_uint4(0, 0, 0, 0);
    if (!kT) {
      const uint32_t row = c / (kBK / 8), kc = (c % (kBK / 8)) * 8u;
      const uint32_t gr = r0 + row, gk = k0 + kc;
      if (gr < rows && gk < k_end)
        v[i] = *reinterpret_cast<const uint4 *>(src + size_t(gr) * ld + gk);
    } else {
      const uint32_t kr = c / (kBM / 8), rc = (c % (kBM / 8)) * 8u;
      const uint32_t gk = k0 + kr, gr = r0 + rc;
      if (gk < k_end && gr < rows)
        v[i] = *reinterpret_cast<const uint4 *>(src + size_t(gk) * ld + gr);
    }
  }
}

template <bool kT>
__device__ __forceinline__ void store_tile(const uint4 (&v)[kChunks], uint16_t *__restrict__ dst) {
#pragma unroll
  for (uint32_t i = 0; i < kChunks; ++i) {
    const uint32_t c = threadIdx.x + 256u * i;
    if (!kT) {
      const uint32_t row = c / (kBK / 8), kc = (c % (kBK / 8)) * 8u;
      *reinterpret_cast<uint4 *>(dst + row * kLdsStride + kc) = v[i];
    } else {
      const uint32_t kr = c / (kBM / 8), rc = (c % (kBM / 8)) * 8u;
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&v[i]);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        dst[(rc + j) * kLdsStride + kr] = e[j];
    }
  }
}

// XCD-aware tile order. Workgroups go to the 8 XCDs round-robin by linear
// id, and each XCD has its own L2. The tiles that share an operand tile (the
// N tiles of one row block, or all tiles of one K split) form a group; the
// members of a group get consecutive slots on one XCD, so the shared operand
// is fetched into one L2 once. Identity when the groups do not fill the XCDs
// evenly.
struct TileId {
  uint32_t x, y, z;
};

__device__ __forceinline__ TileId gemm_tile() {
  const uint32_t gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const uint32_t b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const uint32_t group = gz > 1 ? gx * gy : gy;
  const uint32_t n_groups = gx * gy * gz / group;
  if (n_groups % 8u != 0)
    return {blockIdx.x, blockIdx.y, blockIdx.z};
  const uint32_t xcd = b % 8u, slot = b / 8u;
  const uint32_t gi = (slot / group) * 8u + xcd, member = slot % group;
  if (gz > 1)
    return {member % gx, member / gx, gi};
  return {gi, member, 0u};
}

template <bool kAT, bool kBT, int kEpi>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kBM * kLdsStride];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kBN * kLdsStride];
  const TileId tile = gemm_tile();
  const uint32_t m0 = tile.x * kBM, n0 = tile.y * kBN;
  const uint32_t k_begin = tile.z * g.k_per_split;
  const uint32_t k_end = min(g.K, k_begin + g.k_per_split);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t wm = wave >> 1, wn = wave & 1u;
  const uint16_t *A = reinterpret_cast<const uint16_t *>(g.A);
  const uint16_t *B = reinterpret_cast<const uint16_t *>(g.B);
  f32x16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        acc[i][j][r] = 0.f;
  uint4 va[kChunks], vb[kChunks];
  if (k_begin < k_end) {
    fetch_tile<kAT>(A, g.lda, m0, g.M, k_begin, k_end, va);
    fetch_tile<kBT>(B, g.ldb, n0, g.N, k_begin, k_end, vb);
  }
  for (uint32_t k0 = k_begin; k0 < k_end; k0 += kBK) {
    store_tile<kAT>(va, As);
    store_tile<kBT>(vb, Bs);
    __syncthreads();
    if (k0 + kBK < k_end) {
      fetch_tile<kAT>(A, g.lda, m0, g.M, k0 + kBK, k_end, va);
      fetch_tile<kBT>(B, g.ldb, n0, g.N, k0 + kBK, k_end, vb);
    }
#pragma unroll
    for (uint32_t kk = 0; kk < kBK; kk += 16) {
      bf16x8_t a[2], b[2];
      const uint32_t ko = kk + (lane >> 5) * 8u;
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *reinterpret_cast<const bf16x8_t *>(
            As + (wm * 64u + i * 32u + (lane & 31u)) * kLdsStride + ko);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *reinterpret_cast<const bf16x8_t *>(
            Bs + (wn * 64u + j * 32u + (lane & 31u)) * kLdsStride + ko);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  if constexpr (kEpi == kEpiReluBf16 || kEpi == kEpiMaskBf16) {
    // bf16 results (ReLU, or the mask of the layer input), then the row-major
    // store and, when asked, the transposed one (4 consecutive rows per store)
    uint16_t e[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t row0 = m0 + wm * 64u + i * 32u + q * 8u + (lane >> 5) * 4u;
          const uint32_t col = n0 + wn * 64u + j * 32u + (lane & 31u);
          const bool in = row0 < g.M && col < g.N;
          uint2 mk = make_uint2(0u, 0u);
          if constexpr (kEpi == kEpiMaskBf16) {
            if (in)
              mk = *reinterpret_cast<const uint2 *>(
                  reinterpret_cast<const uint16_t *>(g.mask_t) + size_t(col) * g.ldmt + row0);
          }
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const float v = acc[i][j][4 * q + t];
            float keep;
            if constexpr (kEpi == kEpiMaskBf16) {
              const uint32_t word = t < 2 ? mk.x : mk.y;
              const uint16_t m = uint16_t(word >> (16 * (t & 1)));
              keep = bf2f(m) > 0.f ? v : 0.f;
            } else {
              keep = v > 0.f ? v : 0.f;
            }
            e[i][j][4 * q + t] = f2bf(keep);
            if (in)
              reinterpret_cast<uint16_t *>(g.C)[size_t(row0 + t) * g.ldc + col] =
                  e[i][j][4 * q + t];
          }
        }
    if (g.CT) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t row0 = m0 + wm * 64u + i * 32u + q * 8u + (lane >> 5) * 4u;
            const uint32_t col = n0 + wn * 64u + j * 32u + (lane & 31u);
            if (row0 >= g.M || col >= g.N)
              continue;
            const uint2 v = make_uint2(uint32_t(e[i][j][4 * q]) | (uint32_t(e[i][j][4 * q + 1]) << 16),
                                       uint32_t(e[i][j][4 * q + 2]) | (uint32_t(e[i][j][4 * q + 3]) << 16));
            *reinterpret_cast<uint2 *>(reinterpret_cast<uint16_t *>(g.CT) + size_t(col) * g.ldct +
                                       row0) = v;
          }
    }
  } else {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const uint32_t row = m0 + wm * 64u + i * 32u + (r / 4) * 8u + (lane >> 5) * 4u + (r % 4);
          const uint32_t col = n0 + wn * 64u + j * 32u + (lane & 31u);
          if (row >= g.M || col >= g.N)
            continue;
          const float v = acc[i][j][r];
          const size_t o = size_t(row) * g.ldc + col;
          if constexpr (kEpi == kEpiF32)
            reinterpret_cast<float *>(g.C)[o] = v;
          else
            reinterpret_cast<float *>(g.C)[size_t(tile.z) * g.M * g.N + o] = v;
        }
  }
}

template <bool kAT, bool kBT>
static void launch_gemm_t(const GemmArgs &g, int epi, dim3 grid, hipStream_t s) {
  switch (epi) {
  case kEpiReluBf16:
    hipLaunchKernelGGL((k_gemm<kAT, kBT, kEpiReluBf16>), grid, dim3(256), 0, s, g);
    break;
  case kEpiF32:
    hipLaunchKernelGGL((k_gemm<kAT, kBT, kEpiF32>), grid, dim3(256), 0, s, g);
    break;
  case kEpiMaskBf16:
    hipLaunchKernelGGL((k_gemm<kAT, kBT, kEpiMaskBf16>), grid, dim3(256), 0, s, g);
    break;
  default:
    hipLaunchKernelGGL((k_gemm<kAT, kBT, kEpiSplitF32>), grid, dim3(256), 0, s, g);
    break;
  }
}

hipError_t launch_gemm(const GemmArgs &g, bool a_trans, bool b_trans, int epilogue,
                       uint32_t splits, hipStream_t s) {
  if (g.M % 8 || g.N % 8 || g.K % 8 || g.lda % 8 || g.ldb % 8 || g.k_per_split % 32 ||
      splits == 0 || uint64_t(splits) * g.k_per_split < g.K || (g.CT && g.ldct % 4))
    return hipErrorInvalidValue;
  if (epilogue == kEpiMaskBf16 && (!g.mask_t || g.ldmt % 4))
    return hipErrorInvalidValue;
  dim3 grid((g.M + kBM - 1) / kBM, (g.N + kBN - 1) / kBN, splits);
  if (!a_trans && !b_trans)
    launch_gemm_t<false, false>(g, epilogue, grid, s);
  else if (!a_trans && b_trans)
    launch_gemm_t<false, true>(g, epilogue, grid, s);
  else if (a_trans && b_trans)
    launch_gemm_t<true, true>(g, epilogue, grid, s);
  else
    launch_gemm_t<true, false>(g, epilogue, grid, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Encoding: column c of row i is feature c (c < 9), then for feature d and
// frequency k: sin(pi 2^k x_d), cos(pi 2^k x_d) at 9 + 24 d + 2 k (+1), zero
// from 225 on.
// ---------------------------------------------------------------------------
// The 25 columns of feature d with value x in one encoded row (row[0] is
// column 0): the raw value and, per frequency k, sin and cos(pi 2^k x).
// Frequencies 0, 4 and 8 come from sincospi; the three after each are
// doubled (sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin^2 a), which adds at
// most ~1e-6 absolute error, below the bf16 rounding of the result.
__device__ __forceinline__ void encode_feature(float x, uint32_t d, uint16_t *row) {
  row[d] = f2bf(x);
  uint16_t *o = row + kFeatures + 2 * kFrequencies * d;
  float s = 0.f, c = 1.f;
#pragma unroll
  for (uint32_t k = 0; k < kFrequencies; ++k) {
    if (k % 4 == 0) {
      sincospif(ldexpf(x, int(k)), &s, &c);
    } else {
      const float s2 = 2.f * s * c;
      c = fmaf(-2.f * s, s, 1.f);
      s = s2;
    }
    o[2 * k] = f2bf(s);
    o[2 * k + 1] = f2bf(c);
  }
}

// 28 rows per block: thread t < 252 encodes feature t % 9 of row t / 9, then
// the block zeroes the padding columns of its rows.
constexpr uint32_t kEncRows = 256 / kFeatures;

__global__ __launch_bounds__(256) void k_encode(const float *__restrict__ f, uint32_t n,
                                                uint16_t *__restrict__ out) {
  const size_t row0 = size_t(blockIdx.x) * kEncRows;
  const uint32_t t = threadIdx.x;
  if (t < kEncRows * kFeatures) {
    const size_t row = row0 + t / kFeatures;
    if (row < n)
      encode_feature(f[row * kFeatures + t % kFeatures], t % kFeatures, out + row * kInPad);
  }
  for (uint32_t i = t; i < kEncRows * (kInPad - kEncoded); i += 256u) {
    const size_t row = row0 + i / (kInPad - kEncoded);
    if (row < n)
      out[row * kInPad + kEncoded + i % (kInPad - kEncoded)] = 0;
  }
}

hipError_t launch_encode(const float *features, uint32_t n, bf16 *out, hipStream_t s) {
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_encode, dim3((n + kEncRows - 1) / kEncRows), dim3(256), 0, s, features,
                     n, reinterpret_cast<uint16_t *>(out));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused inference forward: one workgroup (4 waves) per 128-row tile. The
// tile's activations stay in LDS (128 x 256 bf16, row stride 264) from the
// encoding through the six hidden layers; only the 9 input features and the
// 16 outputs per row touch HBM, and the weights (128 KB per layer, shared by
// every tile) come from L2. Wave w computes output columns [64 w, 64 w + 64)
// of all 128 rows (4 x 2 tiles of v_mfma_f32_32x32x16_bf16, weights as the
// A operand so each lane ends with 4 adjacent columns of one row: one 8-byte
// LDS store) with its weight fragments loaded kWAhead K steps ahead; the
// output layer (16 columns) gives wave w rows [32 w, 32 w + 32). The MFMAs
// of every output element run in the same K order as k_gemm's and the
// epilogues round the same way, so the result equals the layer-by-layer
// forward bit for bit (GPU test).
// ---------------------------------------------------------------------------
static_assert(kInPad == kWidth, "hidden weight matrices are all kWidth x kWidth");
constexpr uint32_t kFuseRows = 128, kXStride = kWidth + 8;
// weight fragments in flight: 1 / 2 / 3 / 4 steps ahead measured 1.12 / 1.04 /
// 1.01 / 1.03 ms per 2^20 queries
constexpr uint32_t kWAhead = 3;
constexpr size_t kFuseLds = size_t(kFuseRows) * kXStride * sizeof(uint16_t);

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ feat,
                                                     const uint16_t *__restrict__ w,
                                                     float *__restrict__ pred) {
  extern __shared__ __attribute__((aligned(16))) uint16_t X[];
  const size_t row0 = size_t(blockIdx.x) * kFuseRows;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t l32 = lane & 31u, khalf = (lane >> 5) * 8u;
  for (uint32_t i = threadIdx.x; i < kFuseRows * kFeatures; i += 256u) {
    const uint32_t r = i / kFeatures, d = i % kFeatures;
    encode_feature(feat[(row0 + r) * kFeatures + d], d, X + r * kXStride);
  }
  for (uint32_t i = threadIdx.x; i < kFuseRows * (kInPad - kEncoded); i += 256u)
    X[(i / (kInPad - kEncoded)) * kXStride + kEncoded + i % (kInPad - kEncoded)] = 0;
  __syncthreads();
#pragma unroll 1
  for (uint32_t l = 0; l < kHidden; ++l) {
    // hidden layers are 256 x 256 (layer 0's input is padded to 256)
    const uint16_t *W = w + size_t(l) * kWidth * kWidth;
    f32x16_t acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r)
          acc[i][j][r] = 0.f;
    const uint16_t *wrow[2] = {W + size_t(wave * 64u + l32) * kWidth + khalf,
                               W + size_t(wave * 64u + 32u + l32) * kWidth + khalf};
    // Weight fragments are loaded kWAhead K steps before their MFMAs (a ring
    // of registers); scheduling barriers keep the compiler from sinking the
    // loads next to their use, which left every step waiting on L2. Reading
    // the activation fragments one K step ahead as well measured slower
    // (0.96 vs 0.88 ms per 2^20 queries; 82 instead of 58 VGPRs).
    constexpr uint32_t kSteps = kWidth / 16;
    bf16x8_t bq[kWAhead + 1][2];
#pragma unroll
    for (uint32_t s = 0; s < kWAhead; ++s)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bq[s][j] = *reinterpret_cast<const bf16x8_t *>(wrow[j] + s * 16u);
#pragma unroll
    for (uint32_t ks = 0; ks < kSteps; ++ks) {
      const uint32_t k0 = ks * 16u;
      if (ks + kWAhead < kSteps) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
          bq[(ks + kWAhead) % (kWAhead + 1)][j] =
              *reinterpret_cast<const bf16x8_t *>(wrow[j] + k0 + kWAhead * 16u);
      }
      __builtin_amdgcn_sched_barrier(0);
      bf16x8_t a[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[i] = *reinterpret_cast<const bf16x8_t *>(X + (i * 32u + l32) * kXStride + k0 + khalf);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              bq[ks % (kWAhead + 1)][j], a[i], acc[i][j], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads(); // every wave has read this layer's input
    // acc[i][j] holds the transposed tile: register r of lane l is output
    // column 64 w + 32 j + 8 (r / 4) + 4 (l / 32) + r % 4 of row 32 i + l % 32,
    // so each lane stores 4 adjacent columns at once
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t row = i * 32u + l32;
          const uint32_t col = wave * 64u + j * 32u + q * 8u + (lane >> 5) * 4u;
          bf16x4_t v;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float x = acc[i][j][4 * q + e];
            v[e] = __bf16(x > 0.f ? x : 0.f);
          }
          *reinterpret_cast<bf16x4_t *>(X + row * kXStride + col) = v;
        }
    __syncthreads();
  }
  // output layer: [16 x 256] weights, columns 16..31 of the MFMA tile are zero
  const uint16_t *W = w + size_t(kHidden) * kWidth * kWidth;
  f32x16_t acc;
#pragma unroll
  for (int r = 0; r < 16; ++r)
    acc[r] = 0.f;
  for (uint32_t k0 = 0; k0 < kWidth; k0 += 16) {
    bf16x8_t b;
    if (l32 < kOutPad) {
      b = *reinterpret_cast<const bf16x8_t *>(W + size_t(l32) * kWidth + k0 + khalf);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        b[e] = __bf16(0.f);
    }
    const bf16x8_t a =
        *reinterpret_cast<const bf16x8_t *>(X + (wave * 32u + l32) * kXStride + k0 + khalf);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  if (l32 < kOutPad) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const uint32_t row = wave * 32u + (r / 4) * 8u + (lane >> 5) * 4u + (r % 4);
      pred[(row0 + row) * kOutPad + l32] = acc[r];
    }
  }
}

hipError_t launch_mlp_forward(const float *features, uint32_t rows_pad, const bf16 *w,
                              float *pred, hipStream_t s) {
  if (rows_pad % kFuseRows)
    return hipErrorInvalidValue;
  if (rows_pad == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, dim3(rows_pad / kFuseRows), dim3(256), kFuseLds, s,
                     features, reinterpret_cast<const uint16_t *>(w), pred);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Loss.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ pred,
                                              const float *__restrict__ target, uint32_t n,
                                              uint32_t rows_pad, uint16_t *__restrict__ d_pred,
                                              uint16_t *__restrict__ d_pred_t, uint32_t ld_t,
                                              float *__restrict__ partials) {
  __shared__ float red[256];
  const uint32_t row = blockIdx.x * 256u + threadIdx.x;
  float l = 0.f;
  if (row >= n && row < rows_pad) {
#pragma unroll
    for (uint32_t c = 0; c < kOutPad; ++c) {
      d_pred[size_t(row) * kOutPad + c] = 0;
      if (d_pred_t)
        d_pred_t[size_t(c) * ld_t + row] = 0;
    }
  }
  if (row < n) {
    const float inv = 1.f / (3.f * float(n));
#pragma unroll
    for (uint32_t c = 0; c < kOutPad; ++c) {
      uint16_t d = 0;
      if (c < kOut) {
        const float p = pred[size_t(row) * kOutPad + c], t = target[size_t(row) * kOut + c];
        const float den = p * p + 0.01f;
        const float e = p - t;
        l += e * e / den;
        d = f2bf(2.f * e / den * inv);
      }
      d_pred[size_t(row) * kOutPad + c] = d;
      if (d_pred_t)
        d_pred_t[size_t(c) * ld_t + row] = d;
    }
    l *= inv;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (uint32_t off = 128; off >= 1; off >>= 1) {
    if (threadIdx.x < off)
      red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partials[blockIdx.x] = red[0];
}

hipError_t launch_loss(const float *pred, const float *target, uint32_t n, uint32_t rows_pad,
                       bf16 *d_pred, bf16 *d_pred_t, uint32_t ld_t, float *loss_partials,
                       uint32_t *n_partials, hipStream_t s) {
  // partials cover the n valid rows (blocks past them add nothing)
  *n_partials = (n + 255) / 256;
  const uint32_t blocks = (rows_pad + 255) / 256;
  if (rows_pad < n)
    return hipErrorInvalidValue;
  if (blocks == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, s, pred, target, n, rows_pad,
                     reinterpret_cast<uint16_t *>(d_pred),
                     reinterpret_cast<uint16_t *>(d_pred_t), ld_t, loss_partials);
  return hipGetLastError();
}

// One block: thread t sums partials t, t + 256, ... then a fixed-order tree.
__global__ __launch_bounds__(256) void k_loss_sum(const float *__restrict__ p, uint32_t n,
                                                  float *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += 256u)
    s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t off = 128; off >= 1; off >>= 1) {
    if (threadIdx.x < off)
      red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *out = float(red[0]);
}

hipError_t launch_loss_sum(const float *partials, uint32_t n, float *out, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, partials, n, out);
  return hipGetLastError();
}

// 64 x 64 tiles through LDS (row stride 65 elements).
__global__ __launch_bounds__(256) void k_transpose(const uint16_t *__restrict__ src,
                                                   uint32_t rows, uint32_t cols, uint32_t ld_src,
                                                   uint16_t *__restrict__ dst, uint32_t ld_dst) {
  __shared__ uint16_t tile[64 * 65];
  const uint32_t r0 = blockIdx.y * 64u, c0 = blockIdx.x * 64u;
  for (uint32_t k = threadIdx.x; k < 64u * 64u; k += 256u) {
    const uint32_t r = k / 64u, c = k % 64u;
    if (r0 + r < rows && c0 + c < cols)
      tile[r * 65u + c] = src[size_t(r0 + r) * ld_src + c0 + c];
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < 64u * 64u; k += 256u) {
    const uint32_t c = k / 64u, r = k % 64u;
    if (r0 + r < rows && c0 + c < cols)
      dst[size_t(c0 + c) * ld_dst + r0 + r] = tile[r * 65u + c];
  }
}

hipError_t launch_transpose(const bf16 *src, uint32_t rows, uint32_t cols, uint32_t ld_src,
                            bf16 *dst, uint32_t ld_dst, hipStream_t s) {
  if (rows == 0 || cols == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_transpose, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(256), 0, s,
                     reinterpret_cast<const uint16_t *>(src), rows, cols, ld_src,
                     reinterpret_cast<uint16_t *>(dst), ld_dst);
  return hipGetLastError();
}

// blockIdx.z = layer: the [out][in] matrix at layer l's offset becomes
// [in][out] (every layer but the last is kWidth x kWidth).
__global__ __launch_bounds__(256) void k_transpose_weights(const uint16_t *__restrict__ w,
                                                           uint16_t *__restrict__ w_t) {
  __shared__ uint16_t tile[64 * 65];
  const uint32_t l = blockIdx.z;
  const uint32_t in = kWidth, out = l + 1 == kLayers ? kOutPad : kWidth;
  const size_t off = size_t(l) * kWidth * kWidth;
  const uint32_t r0 = blockIdx.y * 64u, c0 = blockIdx.x * 64u;
  if (r0 >= out)
    return;
  for (uint32_t k = threadIdx.x; k < 64u * 64u; k += 256u) {
    const uint32_t r = k / 64u, c = k % 64u;
    if (r0 + r < out)
      tile[r * 65u + c] = w[off + size_t(r0 + r) * in + c0 + c];
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < 64u * 64u; k += 256u) {
    const uint32_t c = k / 64u, r = k % 64u;
    if (r0 + r < out)
      w_t[off + size_t(c0 + c) * out + r0 + r] = tile[r * 65u + c];
  }
}

hipError_t launch_transpose_weights(const bf16 *w, bf16 *w_t, hipStream_t s) {
  static_assert(kInPad == kWidth, "layer offsets assume kWidth-wide inputs");
  hipLaunchKernelGGL(k_transpose_weights, dim3(kWidth / 64, kWidth / 64, kLayers), dim3(256), 0,
                     s, reinterpret_cast<const uint16_t *>(w), reinterpret_cast<uint16_t *>(w_t));
  return hipGetLastError();
}

// Block of 256 threads per 256 outputs: wave w sums splits w, w + 4, ...
// for its lane's 4 outputs (float4), then the 4 waves' sums are added in
// wave order: a fixed order for every output.
__global__ __launch_bounds__(256) void k_sum_splits(const float *__restrict__ p,
                                                    uint32_t splits, uint64_t count,
                                                    float *__restrict__ out) {
  __shared__ float4 part[4][64];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint64_t i4 = uint64_t(blockIdx.x) * 64u + lane; // float4 index
  const uint64_t n4 = count / 4;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i4 < n4) {
    const float4 *p4 = reinterpret_cast<const float4 *>(p);
#pragma unroll 4
    for (uint32_t z = w; z < splits; z += 4u) {
      const float4 v = p4[uint64_t(z) * n4 + i4];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0 && i4 < n4) {
    float4 s = part[0][lane];
    for (uint32_t k = 1; k < 4; ++k) {
      s.x += part[k][lane].x;
      s.y += part[k][lane].y;
      s.z += part[k][lane].z;
      s.w += part[k][lane].w;
    }
    reinterpret_cast<float4 *>(out)[i4] = s;
  }
}

hipError_t launch_sum_splits(const float *partials, uint32_t splits, uint64_t count,
                             float *out, hipStream_t s) {
  if (count % 4)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sum_splits, dim3(uint32_t((count / 4 + 63) / 64)), dim3(256), 0, s,
                     partials, splits, count, out);
  return hipGetLastError();
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ wb, uint64_t count,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

hipError_t launch_adam(float *w, const float *g, float *m, float *v, bf16 *w_bf16,
                       uint64_t count, float lr, float beta1, float beta2, float eps,
                       uint32_t t, hipStream_t s) {
  const float c1 = 1.f - powf(beta1, float(t)), c2 = 1.f - powf(beta2, float(t));
  hipLaunchKernelGGL(k_adam, dim3(uint32_t((count + 255) / 256)), dim3(256), 0, s, w, g, m, v,
                     reinterpret_cast<uint16_t *>(w_bf16), count, lr, beta1, beta2, eps, c1,
                     c2);
  return hipGetLastError();
}

__global__ void k_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out,
                          uint64_t count) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count)
    out[i] = f2bf(in[i]);
}

hipError_t launch_to_bf16(const float *in, bf16 *out, uint64_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(uint32_t((count + 255) / 256)), dim3(256), 0, s, in,
                     reinterpret_cast<uint16_t *>(out), count);
  return hipGetLastError();
}

__global__ void k_compact_pred(const float *__restrict__ pred, uint32_t n,
                               float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n)
    return;
  out[i] = pred[size_t(i / 3) * kOutPad + i % 3];
}

hipError_t launch_compact_pred(const float *pred, uint32_t n, float *out, hipStream_t s) {
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(k_compact_pred, dim3((3 * n + 255) / 256), dim3(256), 0, s, pred, n, out);
  return hipGetLastError();
}

__global__ void k_gather(const float *__restrict__ feat, const float *__restrict__ targ,
                         const uint32_t *__restrict__ idx, uint32_t n,
                         float *__restrict__ fo, float *__restrict__ to) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const size_t r = idx[i];
#pragma unroll
  for (uint32_t c = 0; c < kFeatures; ++c)
    fo[size_t(i) * kFeatures + c] = feat[r * kFeatures + c];
#pragma unroll
  for (uint32_t c = 0; c < kOut; ++c)
    to[size_t(i) * kOut + c] = targ[r * kOut + c];
}

hipError_t launch_gather(const float *feat, const float *targ, const uint32_t *idx,
                         uint32_t n, float *feat_out, float *targ_out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, feat, targ, idx, n,
                     feat_out, targ_out);
  return hipGetLastError();
}

__global__ void k_random_rows(uint32_t n, uint64_t n_rows, uint64_t seed, uint32_t *idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t(i) + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  idx[i] = uint32_t(z % n_rows);
}

hipError_t launch_random_rows(uint32_t n, uint64_t n_rows, uint64_t seed, uint32_t *idx,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_random_rows, dim3((n + 255) / 256), dim3(256), 0, s, n, n_rows, seed,
                     idx);
  return hipGetLastError();
}

} // namespace mlp
} // namespace pt
