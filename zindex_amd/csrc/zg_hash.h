// zg_hash.h -- hash and seed primitives shared by the gfx950 kernels and the
// host-side index builder. Every function states the reference code it
// computes; results are bit-identical to it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zg {

// utils/ocm-minimizers/src/lib.rs:40-46
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// utils/ocm-minimizers/src/lib.rs:35-38
__host__ __device__ inline uint64_t mixsplit(uint64_t key, uint64_t seed) {
  return splitmix64(key + seed);
}

// The inverse of an odd number modulo 2^64 (Newton steps from a, which is
// its own inverse modulo 8; each step doubles the correct low bits).
__host__ __device__ constexpr uint64_t inverse_odd(uint64_t a) {
  uint64_t x = a;
  for (int i = 0; i < 5; i++) x *= 2 - a * x;
  return x;
}

// splitmix64 run backwards: each xorshift and odd multiply is a bijection of
// 64-bit words, undone in reverse order. mixsplit(key, seed) = h gives
// key = unsplitmix64(h) - seed.
__host__ __device__ inline uint64_t unsplitmix64(uint64_t z) {
  z ^= (z >> 31) ^ (z >> 62);
  z *= inverse_odd(0x94D049BB133111EBULL);
  z ^= (z >> 27) ^ (z >> 54);
  z *= inverse_odd(0xBF58476D1CE4E5B9ULL);
  z ^= (z >> 30) ^ (z >> 60);
  return z - 0x9E3779B97F4A7C15ULL;
}

// storage.rs:3214-3216
__host__ __device__ inline uint64_t fingerprint(uint64_t key, uint64_t seed, uint32_t bits) {
  return splitmix64(key ^ seed ^ 0x8E2D4F6A9B13C5D7ULL) & ((1ULL << bits) - 1);
}

// abandoned.rs:8-10
__host__ __device__ inline uint32_t abandoned_hash(uint64_t key, uint64_t seed) {
  return (uint32_t)(splitmix64(key ^ seed ^ 0xD6E8FEB86659FD93ULL) >> 32);
}

// storage.rs:99-101
__host__ __device__ inline uint64_t partition_id(uint64_t key, uint64_t seed,
                                                 uint64_t partitions) {
  return splitmix64(key ^ seed ^ 0x34B61D8F29A7C053ULL) & (partitions - 1);
}

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Per-bank constants of PositionHash::positions (storage.rs:3468-3480):
// segment length L and the number of admissible first segments.
struct SegmentGeometry {
  uint64_t length;
  uint64_t count;
};

__host__ __device__ inline SegmentGeometry segment_geometry(uint64_t cells, uint32_t hashes,
                                                            uint64_t fixed_length) {
  uint64_t length = fixed_length;
  if (length == 0) {
    uint64_t n = cells / hashes;
    n = n < 1 ? 1 : (n > 4096 ? 4096 : n);
    length = 1;
    while (length * 2 <= n) length *= 2;
  }
  return {length, cells / length - hashes + 1};
}

// PositionHash::new (storage.rs:3458-3465) + positions (3468-3480): slot i.
__host__ __device__ inline uint64_t slot_base(uint64_t first, SegmentGeometry g) {
  return mulhi64(first, g.count) * g.length;
}
__host__ __device__ inline uint64_t slot_at(uint64_t base, uint64_t h, uint32_t i,
                                            SegmentGeometry g) {
  return base + (uint64_t)i * g.length + ((h ^ (h >> 33)) & (g.length - 1));
}

// utils/ocm-minimizers/src/lib.rs:25-34
__host__ __device__ inline uint64_t canonical_encoded(uint64_t encoded, int k) {
  uint64_t value = encoded, rc = 0;
  for (int i = 0; i < k; i++) {
    rc = (rc << 2) | ((value & 3) ^ 3);
    value >>= 2;
  }
  return encoded < rc ? encoded : rc;
}

// Reverses the order of the 32 two-bit groups of x (clang builtin on host and
// device; the second step restores the bit order inside each group).
__host__ __device__ inline uint64_t reverse_pairs(uint64_t x) {
  uint64_t r = __builtin_bitreverse64(x);
  return ((r >> 1) & 0x5555555555555555ULL) | ((r & 0x5555555555555555ULL) << 1);
}

// Canonical k-mer (k <= 32) from 64 bits of packed bases, base p+i at bits
// 2i..2i+1 (A=0,C=1,G=2,T=3): min(forward, reverse complement) exactly as the
// rolling scan_canonical_kmers (src/lib.rs:645-681) produces at position p.
__host__ __device__ inline uint64_t canonical_from_window(uint64_t w, int k) {
  uint64_t mask = k == 32 ? ~0ULL : ((1ULL << (2 * k)) - 1);
  uint64_t rc = ~w & mask;
  uint64_t fwd = reverse_pairs(w) >> (64 - 2 * k);
  return fwd < rc ? fwd : rc;
}

// min(x, reverse complement of x) over w <= 32 bases.
__host__ __device__ inline uint64_t canonical_weight(uint64_t x, uint32_t w) {
  uint64_t mask = w >= 32 ? ~0ULL : ((1ULL << (2 * w)) - 1);
  uint64_t rc = reverse_pairs(~x & mask) >> (64 - 2 * w);
  return x < rc ? x : rc;
}

// A spaced seed (SpacedSeed, src/anchors.rs:10-90): the retained bases of a
// key in their order (pext with the pattern's two-bit mask) taken as runs of
// consecutive retained bases, then made canonical over the seed's weight.
// Runs count bases from the least significant end of the key.
struct SpacedRuns {
  uint32_t n = 0;  // runs; 0 = no pattern
  uint32_t weight = 0;
  uint8_t src[16] = {}, dst[16] = {}, len[16] = {};
};
__host__ __device__ inline uint64_t spaced_canonical(uint64_t key, const SpacedRuns &s) {
  uint64_t r = 0;
  for (uint32_t i = 0; i < s.n; i++) {
    uint64_t m = s.len[i] >= 32 ? ~0ULL : ((1ULL << (2 * s.len[i])) - 1);
    r |= ((key >> (2 * s.src[i])) & m) << (2 * s.dst[i]);
  }
  return canonical_weight(r, s.weight);
}
// The runs of a pattern of `width` bases whose retained base i (from the
// least significant end) has bits 2i..2i+1 of mask set.
inline SpacedRuns spaced_runs(uint32_t width, uint64_t mask) {
  SpacedRuns s;
  uint32_t out = 0;
  for (uint32_t i = 0; i < width;) {
    if (!((mask >> (2 * i)) & 3)) { i++; continue; }
    uint32_t j = i;
    while (j < width && ((mask >> (2 * j)) & 3)) j++;
    s.src[s.n] = (uint8_t)i;
    s.dst[s.n] = (uint8_t)out;
    s.len[s.n] = (uint8_t)(j - i);
    s.n++;
    out += j - i;
    i = j;
  }
  s.weight = out;
  return s;
}

// Randstrobe strobe hash (src/anchors.rs:134-151): XXH64 of one 8-byte
// little-endian input with seed 0, per the published XXH64 specification.
__host__ __device__ inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
__host__ __device__ inline uint64_t strobe_hash(uint64_t v) {
  constexpr uint64_t P1 = 0x9E3779B185EBCA87ULL, P2 = 0xC2B2AE3D27D4EB4FULL,
                     P3 = 0x165667B19E3779F9ULL, P4 = 0x85EBCA77C2B2AE63ULL,
                     P5 = 0x27D4EB2F165667C5ULL;
  uint64_t h = P5 + 8;
  h ^= rotl64(v * P2, 31) * P1;
  h = rotl64(h, 27) * P1 + P4;
  h ^= h >> 33;
  h *= P2;
  h ^= h >> 29;
  h *= P3;
  return h ^ (h >> 32);
}
// The symmetric pair key h1/2 + h2/2 (src/anchors.rs:177-179).
__host__ __device__ inline uint64_t strobe_pair(uint64_t a, uint64_t b) { return a / 2 + b / 2; }
// Partner rank: popcount of the low eight bits of h1 ^ h2 (src/anchors.rs:228-248).
__host__ __device__ inline uint32_t strobe_rank(uint64_t a, uint64_t b) {
  return (uint32_t)__builtin_popcountll((a ^ b) & 255);
}

// Open-closed mod-minimizer parameters (implementation.rs:89-98).
struct OcmParams {
  int w, t, s;
};
__host__ __device__ inline OcmParams ocm_parameters(int k, int m) {
  OcmParams p;
  p.w = k - m + 1;
  int r = m < 4 ? m : 4;
  p.t = r + (m - r) % p.w;
  p.s = p.t < 4 ? p.t : 4;
  if (((p.t - p.s) & 1) != 0) p.s -= 1;
  return p;
}

constexpr uint64_t kOcmInnerSeed = 0xC6BC279692B5CC83ULL;

}  // namespace zg
