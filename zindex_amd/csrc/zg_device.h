// zg_device.h -- small device helpers shared by the gfx950 kernel files:
// wave lane reads, bit-stream reads of the packed reads, read lookup by
// position, wave/block scans, and get_bits over a key's H bank rows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zg_hash.h"
#include "zg_kernels.h"

namespace zg {

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, l);
  uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t readlane32(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// 64 bits of a bit/pair stream starting at bit `bit` (words padded by one).
__device__ __forceinline__ uint64_t stream_bits(const uint64_t *words, uint64_t bit) {
  uint64_t w = bit >> 6;
  uint32_t sh = (uint32_t)(bit & 63);
  uint64_t lo = words[w];
  if (!sh) return lo;
  return (lo >> sh) | (words[w + 1] << (64 - sh));
}

// Largest r in [0, n) with off[r] <= x (off is non-decreasing, off[0] <= x).
__device__ __forceinline__ uint64_t find_segment(const uint64_t *off, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// find_segment by the whole wave: each round its 64 lanes probe 64 evenly
// spaced entries of the range and a ballot keeps the piece holding the
// answer (4 dependent loads at a million segments instead of 20). x and n
// are wave-uniform and every lane is active.
__device__ __forceinline__ uint64_t wave_find_segment(const uint64_t *off, uint64_t n, uint64_t x,
                                                      uint32_t lane) {
  uint64_t lo = 0, hi = n - 1;  // off[lo] <= x; the answer lies in [lo, hi]
  while (lo < hi) {
    const uint64_t step = (hi - lo + 64) / 64;  // 64 pieces of at most step entries
    const uint64_t m = lo + (uint64_t)lane * step;
    const uint64_t ok = __ballot(m <= hi && off[m] <= x);  // lane 0 always
    const uint64_t t = 63 - (uint64_t)__builtin_clzll(ok);
    lo += t * step;
    hi = hi < lo + step - 1 ? hi : lo + step - 1;
  }
  return lo;
}

// Lane src's value of v, src per lane (through the LDS crossbar).
__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)src, 64);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)src, 64);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t lanemask_lt(uint32_t lane) {
  return lane ? (~0ULL >> (64 - lane)) : 0ULL;
}

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, uint32_t d) {
  return (uint64_t)__shfl_up((unsigned long long)v, d, 64);
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_inclusive_scan(uint64_t v, uint32_t lane) {
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint64_t t = shfl_up64(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive prefix sum over a block of NW waves; `total` gets the block sum.
// sh must hold NW words; the call synchronizes the block twice.
template <int NW>
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *sh, uint64_t &total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t inc = wave_inclusive_scan(v, lane);
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  uint64_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) {
    const uint64_t t = sh[w];
    if (w < (int)wave) before += t;
    total += t;
  }
  __syncthreads();
  return before + inc - v;
}

// get_bits (storage.rs:3525-3538) on one row, byte loads (rare paths only).
__device__ inline uint64_t row_bits(const uint8_t *row, uint64_t offset, uint32_t width) {
  uint64_t byte = offset >> 3;
  uint32_t shift = (uint32_t)(offset & 7);
  uint32_t count = (shift + width + 7) >> 3;
  uint64_t v = 0;
  uint32_t first = count < 8 ? count : 8;
  for (uint32_t i = 0; i < first; i++) v |= (uint64_t)row[byte + i] << (8 * i);
  v >>= shift;
  if (count == 9) v |= (uint64_t)row[byte + 8] << (64 - shift);
  if (width < 64) v &= (1ULL << width) - 1;
  return v;
}

// XOR over the key's H rows of get_bits(row, offset, width).
__device__ inline uint64_t xor_rows_bits(const BankDev &bk, uint64_t key, uint64_t seed,
                                  uint64_t offset, uint32_t width) {
  uint64_t first = mixsplit(key, seed);
  uint64_t step = splitmix64(first);
  uint64_t base = slot_base(first, bk.geo);
  uint64_t h = first, v = 0;
  for (uint32_t i = 0; i < bk.hashes; i++) {
    uint64_t slot = slot_at(base, h, i, bk.geo);
    v ^= row_bits(bk.data + slot * bk.row_bytes, offset, width);
    h += step;
  }
  return v;
}

}  // namespace zg
