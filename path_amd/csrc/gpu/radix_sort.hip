// Stable LSD radix sort of (64-bit key, 32-bit value) pairs (radix_sort.h).
#include "radix_sort.h"

namespace pt {
namespace gpu {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 16;
constexpr uint32_t kTile = kThreads * kItems;
constexpr uint32_t kDigits = 256;

struct Pass {
  uint32_t n, n_tiles, shift, mask;
};

__device__ inline uint64_t lanes_below() {
  return (uint64_t(1) << __lane_id()) - 1;
}

// Lanes of this wave whose valid key has the same digit as this lane's
// (only meaningful for valid lanes). Every lane of the wave must call it.
__device__ inline uint64_t same_digit(bool valid, uint32_t d, uint32_t bits) {
  uint64_t m = __ballot(valid);
  for (uint32_t b = 0; b < bits; ++b) {
    const bool set = (d >> b) & 1u;
    const uint64_t on = __ballot(valid && set);
    m &= set ? on : ~on;
  }
  return m;
}

// Exclusive prefix sum of v over the 256 threads of the block; *total gets
// the sum. Every thread must call it; lds holds 4 words.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t x = v;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o)
      x += y;
  }
  if (lane == 63)
    lds[w] = x;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t k = 0; k < kThreads / 64; ++k) {
    const uint32_t c = lds[k];
    before += k < w ? c : 0u;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return before + x - v;
}

__device__ inline uint32_t digit_bits(uint32_t mask) { return 32u - __clz(mask); }

// hist[d * n_tiles + tile] = number of keys of the tile with digit d.
__global__ __launch_bounds__(kThreads) void k_sort_hist(Pass p, const unsigned long long *keys,
                                                        uint32_t *hist) {
  __shared__ uint32_t count[kDigits];
  count[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t bits = digit_bits(p.mask);
  const size_t base = size_t(blockIdx.x) * kTile;
  for (uint32_t r = 0; r < kItems; ++r) {
    const size_t i = base + r * kThreads + threadIdx.x;
    const bool valid = i < p.n;
    const uint32_t d = valid ? uint32_t(keys[i] >> p.shift) & p.mask : 0u;
    const uint64_t peers = same_digit(valid, d, bits);
    if (valid && (peers & lanes_below()) == 0)
      atomicAdd(&count[d], uint32_t(__popcll(peers)));
  }
  __syncthreads();
  hist[size_t(threadIdx.x) * p.n_tiles + blockIdx.x] = count[threadIdx.x];
}

// One block per digit: exclusive scan of the digit's row of hist over tiles;
// totals[d] = keys with digit d.
__global__ __launch_bounds__(kThreads) void k_sort_scan(Pass p, uint32_t *hist,
                                                        uint32_t *totals) {
  __shared__ uint32_t lds[kThreads / 64];
  uint32_t *row = hist + size_t(blockIdx.x) * p.n_tiles;
  uint32_t carry = 0;
  for (uint32_t t0 = 0; t0 < p.n_tiles; t0 += kThreads) {
    const uint32_t t = t0 + threadIdx.x;
    const uint32_t v = t < p.n_tiles ? row[t] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, lds, &tot);
    if (t < p.n_tiles)
      row[t] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0)
    totals[blockIdx.x] = carry;
}

// Stable scatter of one tile: a key's output position is the start of its
// digit (keys of smaller digits), plus the keys of its digit in earlier tiles
// (scanned hist), earlier rounds of this tile, earlier waves of its round and
// lower lanes of its wave.
__global__ __launch_bounds__(kThreads) void k_sort_scatter(Pass p, const unsigned long long *keys,
                                                           const uint32_t *vals,
                                                           const uint32_t *hist,
                                                           const uint32_t *totals,
                                                           unsigned long long *keys_out,
                                                           uint32_t *vals_out) {
  __shared__ uint32_t offs[kDigits];
  __shared__ uint32_t wave_count[kThreads / 64][kDigits];
  __shared__ uint32_t lds[kThreads / 64];
  const uint32_t tid = threadIdx.x, w = tid >> 6;
  uint32_t all;
  const uint32_t start = block_exclusive_scan(totals[tid], lds, &all);
  offs[tid] = start + hist[size_t(tid) * p.n_tiles + blockIdx.x];
  for (uint32_t k = 0; k < kThreads / 64; ++k)
    wave_count[k][tid] = 0;
  __syncthreads();
  const uint32_t bits = digit_bits(p.mask);
  const size_t base = size_t(blockIdx.x) * kTile;
  for (uint32_t r = 0; r < kItems; ++r) {
    const size_t i = base + r * kThreads + tid;
    const bool valid = i < p.n;
    const unsigned long long key = valid ? keys[i] : 0ull;
    const uint32_t d = uint32_t(key >> p.shift) & p.mask;
    const uint64_t peers = same_digit(valid, d, bits);
    const uint64_t below = peers & lanes_below();
    if (valid && below == 0)
      wave_count[w][d] = uint32_t(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t pos = offs[d] + uint32_t(__popcll(below));
      for (uint32_t k = 0; k < w; ++k)
        pos += wave_count[k][d];
      keys_out[pos] = key;
      vals_out[pos] = vals[i];
    }
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t k = 0; k < kThreads / 64; ++k) {
      add += wave_count[k][tid];
      wave_count[k][tid] = 0;
    }
    offs[tid] += add;
    __syncthreads();
  }
}

uint32_t tiles_of(uint32_t n) { return (n + kTile - 1) / kTile; }

} // namespace

size_t radix_sort_temp_bytes(uint32_t n) {
  return (size_t(kDigits) * tiles_of(n) + kDigits) * sizeof(uint32_t);
}

hipError_t radix_sort_pairs(unsigned long long *keys_in, uint32_t *vals_in,
                            unsigned long long *keys_out, uint32_t *vals_out, uint32_t n,
                            uint32_t end_bit, void *temp, size_t temp_bytes, hipStream_t s) {
  if (end_bit < 1 || end_bit > 64)
    return hipErrorInvalidValue;
  if (n == 0)
    return hipSuccess;
  if (temp_bytes < radix_sort_temp_bytes(n) || !temp)
    return hipErrorInvalidValue;
  const uint32_t n_tiles = tiles_of(n);
  uint32_t *hist = static_cast<uint32_t *>(temp);
  uint32_t *totals = hist + size_t(kDigits) * n_tiles;
  unsigned long long *k[2] = {keys_in, keys_out};
  uint32_t *v[2] = {vals_in, vals_out};
  int cur = 0;
  for (uint32_t shift = 0; shift < end_bit; shift += 8) {
    const uint32_t bits = end_bit - shift < 8 ? end_bit - shift : 8;
    const Pass p{n, n_tiles, shift, (1u << bits) - 1u};
    hipLaunchKernelGGL(k_sort_hist, dim3(n_tiles), dim3(kThreads), 0, s, p, k[cur], hist);
    hipLaunchKernelGGL(k_sort_scan, dim3(kDigits), dim3(kThreads), 0, s, p, hist, totals);
    hipLaunchKernelGGL(k_sort_scatter, dim3(n_tiles), dim3(kThreads), 0, s, p, k[cur], v[cur],
                       hist, totals, k[cur ^ 1], v[cur ^ 1]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    cur ^= 1;
  }
  if (cur == 0) {
    // an even number of passes ends in the input arrays
    hipError_t e = hipMemcpyAsync(keys_out, keys_in, size_t(n) * sizeof(*keys_in),
                                  hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(vals_out, vals_in, size_t(n) * sizeof(*vals_in),
                         hipMemcpyDeviceToDevice, s);
    return e;
  }
  return hipSuccess;
}

} // namespace gpu
} // namespace pt
