// Wave64 primitives on DPP / cross-lane operations (gfx9 DPP controls).
#pragma once

#include <hip/hip_runtime.h>

namespace pt {
namespace gpu {

// DPP control words (gfx9): row_shl:n = 0x100 + n, row_shr:n = 0x110 + n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), kCtrl, kRowMask, 0xf,
                                              false));
}

// Inclusive XOR prefix over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_xor_scan(uint32_t x) {
  x ^= dpp_u32<0x111>(x);
  x ^= dpp_u32<0x112>(x);
  x ^= dpp_u32<0x114>(x);
  x ^= dpp_u32<0x118>(x);
  x ^= dpp_u32<0x142, 0xa>(x); // rows 1, 3 take lane 15 of the row below
  x ^= dpp_u32<0x143, 0xc>(x); // rows 2, 3 take lane 31
  return x;
}

// Value of lane (lane + off), for the halving reduction tree. Within a row of
// 16 lanes this is DPP row_shl; larger offsets use ds_bpermute.
template <int kOff> __device__ __forceinline__ float shfl_down_f(float v) {
  if constexpr (kOff < 16) {
    return __uint_as_float(dpp_u32<0x100 + kOff>(__float_as_uint(v)));
  } else {
    return __shfl_down(v, kOff, 64);
  }
}

// Sum of the values of `group` consecutive lanes (group a power of two <= 64)
// left in the first lane of each group: v[i] += v[i + off] for
// off = group/2, ..., 1. Lanes other than the first of a group hold partial
// values.
__device__ __forceinline__ float group_reduce_sum(float v, uint32_t group) {
  if (group > 32)
    v = v + shfl_down_f<32>(v);
  if (group > 16)
    v = v + shfl_down_f<16>(v);
  if (group > 8)
    v = v + shfl_down_f<8>(v);
  if (group > 4)
    v = v + shfl_down_f<4>(v);
  if (group > 2)
    v = v + shfl_down_f<2>(v);
  if (group > 1)
    v = v + shfl_down_f<1>(v);
  return v;
}

// Exclusive prefix count of set predicates below this lane and the total,
// from one ballot (v_mbcnt).
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t ballot_prefix(bool pred, uint32_t *total) {
  uint64_t mask = __ballot(pred);
  *total = uint32_t(__popcll(mask));
  uint32_t lo = uint32_t(mask), hi = uint32_t(mask >> 32);
  return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

} // namespace gpu
} // namespace pt
