// zg_kernels.hip -- gfx950 kernels of the zorindex query path.
//
// Data layout in HBM (one query batch):
//   codes    2-bit bases, 32 per u64, base p at bits 2(p%32) of word p/32
//   invalid  1 bit per base, 64 per u64; set for non-ACGT bytes, for the one
//            separator position after every read, and for padding
//   bank     slots * row_bytes bytes, plane layout: the G fingerprint fields of
//            a row first (b bits each, little endian), the membership masks of
//            multi-member groups after them
// The lookup is HBM-bound random row gathers: every seed reads H rows of the
// first ceil(G*b/8) bytes. A wave owns a tile of 64 lanes x VB bytes of those
// rows, so one seed costs H coalesced 64*VB-byte reads per tile, XOR and
// compare stay in registers and per-read colour counters live in VGPRs.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "zg_device.h"
#include "zg_kernels.h"

namespace zg {

namespace {

// A/a -> 0, C/c -> 1, G/g -> 2, T/t -> 3, anything else -1, without
// branches: u = c with bit 5 cleared is one of 0x41, 0x43, 0x47, 0x54 for a
// base (offsets 0, 2, 6, 19 from 'A': mask 0x80045), and ((u >> 1) ^ (u >> 2))
// & 3 is its code.
__device__ __forceinline__ int encode_ascii(uint8_t c) {
  const uint32_t u = c & 0xDFu, d = u - 0x41u;
  const bool base = d < 20u && ((0x80045u >> d) & 1u);
  return base ? (int)(((u >> 1) ^ (u >> 2)) & 3u) : -1;
}

}  // namespace

// ---------------------------------------------------------------------------
// Fills
// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u64_kernel(uint64_t *p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 65536) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)(g > cap ? cap : g);
}

void launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
  if (!n) return;
  fill_u32_kernel<<<grid_for(n, 256, 16384), 256, 0, s>>>(p, n, v);
}
void launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s) {
  if (!n) return;
  fill_u64_kernel<<<grid_for(n, 256, 16384), 256, 0, s>>>(p, n, v);
}

// ---------------------------------------------------------------------------
// Packing: one wave per 64 output words. In round i the wave's lanes take the
// 64 consecutive positions of word w0 + i, so each byte load instruction
// reads 64 consecutive bytes of a read; ballots give the word's invalid bits
// and the two bit planes of its codes, and lane i keeps word w0 + i for one
// coalesced store at the end.
// ---------------------------------------------------------------------------
// Bits 0..31 of x to the even bits of the result.
__device__ __forceinline__ uint64_t spread_even(uint64_t x) {
  x &= 0xFFFFFFFFULL;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFULL;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFULL;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FULL;
  x = (x | (x << 2)) & 0x3333333333333333ULL;
  x = (x | (x << 1)) & 0x5555555555555555ULL;
  return x;
}

__global__ __launch_bounds__(256) void pack_kernel(const uint8_t *seq, const uint64_t *read_off,
                                                   const uint64_t *pos_off, uint64_t n_reads,
                                                   uint64_t positions, uint64_t *codes,
                                                   uint64_t *invalid) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_words = (positions + 63) / 64;
  const uint64_t w0 = ((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6) * 64;
  if (w0 >= n_words) return;  // wave-uniform
  const uint64_t w1 = w0 + 64 < n_words ? w0 + 64 : n_words;
  uint64_t p = w0 * 64 + lane;
  // the lane's read: its positions end at next_pos, byte = p + delta while
  // below byte_end (the read's last position is the separator); found from
  // the read holding the wave's first position
  uint64_t r = wave_find_segment(pos_off, n_reads, w0 * 64, lane);
  const uint64_t px = p < positions ? p : positions - 1;
  while (pos_off[r + 1] <= px) r++;  // pos_off[n_reads] = positions > px
  uint64_t next_pos = pos_off[r + 1], delta = read_off[r] - pos_off[r], byte_end = read_off[r + 1];
  uint64_t my_inv = 0, my_b0 = 0, my_b1 = 0;  // lane i: word w0 + i's ballots
  constexpr int U = 16;  // rounds whose byte loads are in flight together
  for (uint64_t w = w0; w < w1; w += U, p += 64 * U) {
    uint64_t at[U];
    bool have[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t pu = p + 64 * (uint64_t)u;
      have[u] = false;
      at[u] = 0;
      if (w + u < w1 && pu < positions) {
        while (pu >= next_pos) {
          r++;
          next_pos = pos_off[r + 1];
          delta = read_off[r] - pos_off[r];
          byte_end = read_off[r + 1];
        }
        at[u] = pu + delta;
        have[u] = at[u] < byte_end;
      }
    }
    uint8_t ch[U];
#pragma unroll
    for (int u = 0; u < U; u++) ch[u] = have[u] ? seq[at[u]] : (uint8_t)'N';
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int code = encode_ascii(ch[u]);
      const uint64_t inv = __ballot(code < 0);
      const uint64_t b0 = __ballot(code > 0 && (code & 1));
      const uint64_t b1 = __ballot(code > 0 && (code & 2));
      const bool mine = lane == (uint32_t)(w + u - w0);
      my_inv = mine ? inv : my_inv;  // invalid positions hold code 0
      my_b0 = mine ? b0 : my_b0;
      my_b1 = mine ? b1 : my_b1;
    }
  }
  if (w0 + lane < w1) {
    invalid[w0 + lane] = my_inv;
    codes[2 * (w0 + lane)] = spread_even(my_b0) | (spread_even(my_b1) << 1);
    codes[2 * (w0 + lane) + 1] = spread_even(my_b0 >> 32) | (spread_even(my_b1 >> 32) << 1);
  }
}

void launch_pack(const uint8_t *seq, const uint64_t *read_off, const uint64_t *pos_off,
                 uint64_t n_reads, uint64_t positions, uint64_t *codes, uint64_t *invalid,
                 hipStream_t s) {
  uint64_t words = (positions + 63) / 64;
  if (!words || !n_reads) return;
  const uint64_t waves = (words + 63) / 64;
  pack_kernel<<<(uint32_t)((waves + 3) / 4), 256, 0, s>>>(seq, read_off, pos_off, n_reads, positions,
                                                          codes, invalid);
}

// ---------------------------------------------------------------------------
// Homopolymer compression on the device (homopolymer_compress,
// src/lib.rs:209-222). A position is kept unless it holds a valid base equal
// to the base before it; invalid bytes and read separators are kept (they
// break runs). Each read's kept bases are packed to the start of its own
// region of the position space; the rest of the region stays invalid, so
// every later kernel sees the compressed read followed by invalid positions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int position_code(const uint8_t *seq, const uint64_t *read_off,
                                             const uint64_t *pos_off, uint64_t r, uint64_t p) {
  const uint64_t off = p - pos_off[r];
  const uint64_t start = read_off[r], len = read_off[r + 1] - start;
  return off < len ? encode_ascii(seq[start + off]) : -1;
}

// keep_bits[w]: the kept positions among 64w..64w+63; counts[w] their number.
__global__ void hpc_keep_kernel(const uint8_t *seq, const uint64_t *read_off,
                                const uint64_t *pos_off, uint64_t n_reads, uint64_t positions,
                                uint64_t *keep_bits, uint32_t *counts) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t n_words = (positions + 63) / 64;
  if (w >= n_words) return;
  const uint64_t p0 = w * 64;
  uint64_t r = find_segment(pos_off, n_reads, p0 ? p0 - 1 : 0);
  int prev = -1;
  if (p0) prev = position_code(seq, read_off, pos_off, r, p0 - 1);
  uint64_t keep = 0;
  for (int i = 0; i < 64; i++) {
    const uint64_t p = p0 + i;
    if (p >= positions) break;
    while (pos_off[r + 1] <= p) r++;
    const int c = position_code(seq, read_off, pos_off, r, p);
    if (c < 0 || c != prev) keep |= 1ULL << i;
    prev = c;
  }
  keep_bits[w] = keep;
  counts[w] = (uint32_t)__popcll(keep);
}

// Writes each kept valid base at its read's start plus the number of kept
// positions before it in the read. codes must be zero and invalid all ones
// beforehand. One thread per 64 positions; writes are merged per target word.
__global__ void hpc_place_kernel(const uint8_t *seq, const uint64_t *read_off,
                                 const uint64_t *pos_off, uint64_t n_reads, uint64_t positions,
                                 const uint64_t *keep_bits, const uint64_t *keep_off,
                                 uint64_t *codes, uint64_t *invalid) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t n_words = (positions + 63) / 64;
  if (w >= n_words) return;
  const uint64_t p0 = w * 64;
  const uint64_t kb = keep_bits[w];
  if (!kb) return;
  uint64_t r = find_segment(pos_off, n_reads, p0);
  auto rank = [&](uint64_t p) {
    const uint64_t below = p & 63;
    return keep_off[p >> 6] + __popcll(keep_bits[p >> 6] & (below ? (~0ULL >> (64 - below)) : 0));
  };
  uint64_t read_rank = rank(pos_off[r]);
  uint64_t cw = ~0ULL, cbits = 0, iw = ~0ULL, ibits = 0;
  for (uint64_t m = kb; m; m &= m - 1) {
    const int i = __builtin_ctzll(m);
    const uint64_t p = p0 + i;
    if (pos_off[r + 1] <= p) {
      while (pos_off[r + 1] <= p) r++;
      read_rank = rank(pos_off[r]);
    }
    const int c = position_code(seq, read_off, pos_off, r, p);
    if (c < 0) continue;
    const uint64_t q = pos_off[r] + (keep_off[w] + __popcll(kb & (i ? (~0ULL >> (64 - i)) : 0)) -
                                     read_rank);
    if ((q >> 5) != cw) {
      if (cbits) atomicOr((unsigned long long *)&codes[cw], (unsigned long long)cbits);
      cw = q >> 5;
      cbits = 0;
    }
    cbits |= (uint64_t)c << (2 * (q & 31));
    if ((q >> 6) != iw) {
      if (ibits) atomicAnd((unsigned long long *)&invalid[iw], (unsigned long long)~ibits);
      iw = q >> 6;
      ibits = 0;
    }
    ibits |= 1ULL << (q & 63);
  }
  if (cbits) atomicOr((unsigned long long *)&codes[cw], (unsigned long long)cbits);
  if (ibits) atomicAnd((unsigned long long *)&invalid[iw], (unsigned long long)~ibits);
}

void launch_hpc_keep(const uint8_t *seq, const uint64_t *read_off, const uint64_t *pos_off,
                     uint64_t n_reads, uint64_t positions, uint64_t *keep_bits,
                     uint32_t *counts, hipStream_t s) {
  uint64_t words = (positions + 63) / 64;
  if (!words || !n_reads) return;
  hpc_keep_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(
      seq, read_off, pos_off, n_reads, positions, keep_bits, counts);
}

void launch_hpc_place(const uint8_t *seq, const uint64_t *read_off, const uint64_t *pos_off,
                      uint64_t n_reads, uint64_t positions, const uint64_t *keep_bits,
                      const uint64_t *keep_off, uint64_t *codes, uint64_t *invalid,
                      hipStream_t s) {
  uint64_t words = (positions + 63) / 64;
  if (!words || !n_reads) return;
  hpc_place_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(
      seq, read_off, pos_off, n_reads, positions, keep_bits, keep_off, codes, invalid);
}

// ---------------------------------------------------------------------------
// Valid k-mer windows per read (k <= 32): one thread per 64 window starts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t invalid_window_starts(const uint64_t *invalid, uint64_t w,
                                                          uint32_t k) {
  // bit i set iff window [64w+i, 64w+i+k) contains an invalid base (the
  // invalid words are padded for the window length, invalid_words)
  uint64_t m = 0;
  if (k <= 64) {
    // over the 128 bits (lo, hi): a span of s bits ORed at each position,
    // doubled while 2s <= k, then one more OR at offset k - s (< s) covers k
    uint64_t lo = invalid[w], hi = invalid[w + 1];
    uint32_t s = 1;
    while (2 * s <= k) {
      lo |= (lo >> s) | (hi << (64 - s));
      hi |= hi >> s;
      s *= 2;
    }
    if (s < k) lo |= (lo >> (k - s)) | (hi << (64 - (k - s)));
    m = lo;
  } else {
    for (uint32_t d = 0; d < k; d++) m |= stream_bits(invalid, 64 * w + d);
  }
  return m;
}

__global__ void count_kmers_kernel(const uint64_t *invalid, const uint64_t *codes, uint64_t seed,
                                   uint64_t sampling, const uint64_t *pos_off,
                                   uint64_t n_reads, uint64_t positions, uint32_t k,
                                   uint64_t *totals) {
  const uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t n_words = (positions + 63) / 64;
  if ((w & ~63ULL) >= n_words) return;  // wave-uniform
  // the read holding the wave's first position, then this word's first
  uint64_t r = wave_find_segment(pos_off, n_reads, (w & ~63ULL) * 64, threadIdx.x & 63);
  if (w >= n_words) return;
  uint64_t valid = ~invalid_window_starts(invalid, w, k);
  uint64_t p0 = w * 64;
  while (pos_off[r + 1] <= p0) r++;  // pos_off[n_reads] = positions > p0
  if (sampling)  // scan_modimizers (src/lib.rs:796-806)
    for (uint64_t m = valid; m; m &= m - 1) {
      const int j = __builtin_ctzll(m);
      const uint64_t key = canonical_from_window(stream_bits(codes, 2 * (p0 + j)), (int)k);
      if (mixsplit(key, seed) % sampling) valid &= ~(1ULL << j);
    }
  while (r < n_reads && pos_off[r] < p0 + 64) {
    uint64_t a = pos_off[r] > p0 ? pos_off[r] - p0 : 0;
    uint64_t b = pos_off[r + 1] - p0;
    if (b > 64) b = 64;
    if (b > a) {
      uint64_t span = b - a;
      uint64_t mask = (span == 64 ? ~0ULL : ((1ULL << span) - 1)) << a;
      uint64_t c = __popcll(valid & mask);
      if (c) atomicAdd((unsigned long long *)&totals[r], (unsigned long long)c);
    }
    if (pos_off[r + 1] >= p0 + 64) break;
    r++;
  }
}

void launch_count_kmers(const uint64_t *invalid, const uint64_t *codes, uint64_t seed,
                        uint64_t sampling, const uint64_t *pos_off, uint64_t n_reads,
                        uint64_t positions, uint32_t k, uint64_t *totals, hipStream_t s) {
  uint64_t words = (positions + 63) / 64;
  if (!words || !n_reads) return;
  count_kmers_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(
      invalid, codes, seed, sampling, pos_off, n_reads, positions, k, totals);
}

// ---------------------------------------------------------------------------
// Open-closed mod-minimizers (utils/ocm-minimizers/src/implementation.rs).
// A block evaluates 256 consecutive window starts plus the one before them.
// Inner s-mer hashes and anchor ranks (class, t-mer hash) for the block's
// positions are staged in LDS; every thread then takes the minimum rank over
// the k-t+1 anchors of its window, keeping the leftmost and the rightmost tie,
// and picks by the GT majority of the window. A window is emitted when it
// starts a run (the previous window is invalid) or its position differs from
// the previous window's, which is the reference's `previous != pos` rule.
// ---------------------------------------------------------------------------
constexpr uint32_t kOcmTile = 256;
constexpr uint32_t kOcmSpan = kOcmTile + 1 + 256;  // window starts - 1 .. + k (k <= 255)

uint32_t ocm_block_positions() { return kOcmTile; }

// An anchor's rank with the leftmost and rightmost local index holding it;
// combining two ranges keeps the smaller (class, hash) and, on a tie, the
// union of their index spans (ScalarRank::combine with first/last ties,
// implementation.rs:266-346).
struct OcmMin {
  uint64_t h;
  uint32_t c, first, last;
};
__device__ __forceinline__ OcmMin ocm_combine(const OcmMin &a, const OcmMin &b) {
  if (a.c < b.c || (a.c == b.c && a.h < b.h)) return a;
  if (b.c < a.c || (b.c == a.c && b.h < a.h)) return b;
  return {a.h, a.c, a.first < b.first ? a.first : b.first, a.last > b.last ? a.last : b.last};
}

// Open-closed mod-minimizer selection for kOcmTile window starts per block.
// Per window in O(1): invalid-base and GT counts come from 64-position ballot
// masks with popcount prefixes; the minimum anchor over the window's
// W = k - t + 1 anchors comes from a sparse table of range minima (two
// overlapping power-of-two ranges cover the window).
__global__ __launch_bounds__(256) void ocm_select_kernel(
    const uint64_t *codes, const uint64_t *invalid, uint64_t positions, uint32_t k,
    uint32_t m, uint64_t seed, const uint64_t *inner_table, uint8_t *flag, uint8_t *delta,
    uint32_t *block_counts) {
  constexpr uint32_t kWords = (kOcmSpan + 63) / 64 + 1;
  __shared__ uint8_t code_s[kOcmSpan];
  __shared__ uint64_t bad_m[kWords], gt_m[kWords];
  __shared__ uint32_t bad_c[kWords], gt_c[kWords];
  __shared__ uint64_t ih_s[kOcmSpan];
  __shared__ uint64_t th[2][kOcmSpan];
  __shared__ uint8_t tc[2][kOcmSpan];
  __shared__ uint16_t tf[2][kOcmSpan], tl[2][kOcmSpan];
  __shared__ uint32_t pos_s[kOcmTile + 1];
  __shared__ uint8_t ok_s[kOcmTile + 1];
  __shared__ uint32_t wave_counts[4];

  const OcmParams P = ocm_parameters((int)k, (int)m);
  const uint64_t S0 = (uint64_t)blockIdx.x * kOcmTile;
  const int64_t base = (int64_t)S0 - 1;  // position of local index 0
  const uint32_t span = kOcmTile + k;      // local positions needed: base .. base+span-1
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t words = (span + 63) / 64;

  // codes, and the invalid / GT bits of 64 positions per ballot (positions
  // outside the batch count as invalid)
  for (uint32_t i0 = tid & ~63u; i0 < words * 64; i0 += blockDim.x) {
    const uint32_t i = i0 + lane;
    const int64_t p = base + (int64_t)i;
    uint8_t c = 0;
    bool bad = true;
    if (i < span && p >= 0 && (uint64_t)p < positions) {
      const uint64_t up = (uint64_t)p;
      c = (uint8_t)((codes[up >> 5] >> (2 * (up & 31))) & 3);
      bad = (invalid[up >> 6] >> (up & 63)) & 1;
    }
    if (i < span) code_s[i] = c;
    const uint64_t bm = __ballot(bad), gm = __ballot(!bad && c >= 2);
    if (lane == 0) {
      bad_m[i0 >> 6] = bm;
      gt_m[i0 >> 6] = gm;
    }
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t b = 0, g = 0;
    for (uint32_t w = 0; w < words; w++) {
      bad_c[w] = b;
      gt_c[w] = g;
      b += (uint32_t)__popcll(bad_m[w]);
      g += (uint32_t)__popcll(gt_m[w]);
    }
    bad_c[words] = b;
    gt_c[words] = g;
    bad_m[words] = gt_m[words] = 0;
  }
  const uint32_t s = (uint32_t)P.s, t = (uint32_t)P.t;
  for (uint32_t i = tid; i + s <= span; i += blockDim.x) {
    uint32_t x = 0;
    for (uint32_t j = 0; j < s; j++) x = (x << 2) | code_s[i + j];
    ih_s[i] = inner_table[x];
  }
  __syncthreads();
  // anchors: rank (class, mixsplit of the canonical t-mer), sparse-table level 0
  const uint32_t n_anchor = span - t + 1;
  for (uint32_t i = tid; i < n_anchor; i += blockDim.x) {
    uint64_t mn = ~0ULL;
    for (uint32_t j = 0; j + s <= t; j++) mn = ih_s[i + j] < mn ? ih_s[i + j] : mn;
    uint8_t cls;
    if (ih_s[i + (t - s) / 2] == mn) cls = 0;
    else if (ih_s[i] == mn || ih_s[i + t - s] == mn) cls = 1;
    else cls = 2;
    // the canonical t-mer straight from the packed codes (anchors reaching
    // past the batch belong to no valid window)
    const int64_t p = base + (int64_t)i;
    const uint64_t canon = p >= 0 && (uint64_t)p + t <= positions
                               ? canonical_from_window(stream_bits(codes, 2 * (uint64_t)p), (int)t)
                               : 0;
    th[0][i] = mixsplit(canon, seed);
    tc[0][i] = cls;
    tf[0][i] = tl[0][i] = (uint16_t)i;
  }
  __syncthreads();
  // range minima over 2^j anchors, j = 1 .. L with 2^L <= W
  const uint32_t W = k - t + 1;
  uint32_t L = 0;
  while ((2u << L) <= W) L++;
  for (uint32_t j = 1; j <= L; j++) {
    const uint32_t src = (j - 1) & 1, dst = j & 1, half = 1u << (j - 1);
    for (uint32_t i = tid; i + (1u << j) <= n_anchor; i += blockDim.x) {
      OcmMin a{th[src][i], tc[src][i], tf[src][i], tl[src][i]};
      OcmMin b{th[src][i + half], tc[src][i + half], tf[src][i + half], tl[src][i + half]};
      OcmMin c = ocm_combine(a, b);
      th[dst][i] = c.h;
      tc[dst][i] = (uint8_t)c.c;
      tf[dst][i] = (uint16_t)c.first;
      tl[dst][i] = (uint16_t)c.last;
    }
    __syncthreads();
  }
  const uint32_t top = L & 1;
  auto ones = [&](const uint64_t *msk, const uint32_t *cum, uint32_t x) {
    return cum[x >> 6] + (uint32_t)__popcll(msk[x >> 6] & ((1ULL << (x & 63)) - 1));
  };
  // windows: local index w = 0 is start base (S0 - 1), w = 1 + tid is S0 + tid
  for (uint32_t w = tid; w < kOcmTile + 1; w += blockDim.x) {
    const int64_t start = base + (int64_t)w;
    bool ok = start >= 0 && (uint64_t)start + k <= positions &&
              ones(bad_m, bad_c, w + k) == ones(bad_m, bad_c, w);
    uint32_t pos = 0;
    if (ok) {
      const uint32_t gt = ones(gt_m, gt_c, w + k) - ones(gt_m, gt_c, w);
      const uint32_t j2 = w + W - (1u << L);
      OcmMin a{th[top][w], tc[top][w], tf[top][w], tl[top][w]};
      OcmMin b{th[top][j2], tc[top][j2], tf[top][j2], tl[top][j2]};
      OcmMin c = ocm_combine(a, b);
      const uint32_t anchor = (gt > k / 2 ? c.last : c.first) - w;
      pos = anchor % (uint32_t)P.w;
    }
    ok_s[w] = ok;
    pos_s[w] = pos;
  }
  __syncthreads();
  uint64_t S = S0 + tid;
  bool emit = false;
  uint8_t d = 0;
  if (S < positions && ok_s[tid + 1]) {
    d = (uint8_t)pos_s[tid + 1];
    // previous window start S-1 selects position (S-1) + pos_s[tid]
    emit = !ok_s[tid] || (uint64_t)(S - 1 + pos_s[tid]) != S + d;
  }
  if (S < positions) {
    flag[S] = emit;
    delta[S] = d;
  }
  uint64_t b = __ballot(emit);
  if ((tid & 63) == 0) wave_counts[tid >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (tid == 0)
    block_counts[blockIdx.x] = wave_counts[0] + wave_counts[1] + wave_counts[2] + wave_counts[3];
}

// Compaction of the selected windows: one wave per tile of kOcmTile window
// starts (the unit the selection counts in), 64 starts per step; an emitted
// window's key goes to the tile's offset plus the emitted windows before it.
// Waves stride over the tiles.
__global__ __launch_bounds__(256) void ocm_compact_kernel(const uint64_t *codes,
                                                          const uint8_t *flag,
                                                          const uint8_t *delta,
                                                          const uint64_t *tile_off,
                                                          uint64_t positions, uint32_t m,
                                                          SpacedRuns spaced, uint64_t *keys,
                                                          uint64_t *key_win) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_tiles = (positions + kOcmTile - 1) / kOcmTile;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
       tile < n_tiles; tile += waves) {
    constexpr uint32_t Q = kOcmTile / 64;
    const uint64_t S0 = tile * kOcmTile;
    // the tile's flags and deltas, then the code words of its selected
    // windows, each group of loads issued together
    uint32_t f[Q], d[Q];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint64_t S = S0 + 64 * q + lane;
      f[q] = S < positions ? flag[S] : 0;
      d[q] = S < positions ? delta[S] : 0;
    }
    uint64_t lo[Q], hi[Q];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint64_t pos = S0 + 64 * q + lane + d[q];
      const uint32_t sh = 2 * (uint32_t)(pos & 31);
      lo[q] = f[q] ? codes[pos >> 5] : 0;
      hi[q] = f[q] && sh ? codes[(pos >> 5) + 1] : 0;
    }
    uint64_t at = tile_off[tile];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
      const uint64_t b = __ballot(f[q] != 0);
      if (f[q]) {
        const uint64_t S = S0 + 64 * q + lane, pos = S + d[q];
        const uint32_t sh = 2 * (uint32_t)(pos & 31);
        const uint64_t i = at + (uint64_t)__popcll(b & ((1ULL << lane) - 1));
        uint64_t key = canonical_from_window(sh ? (lo[q] >> sh) | (hi[q] << (64 - sh)) : lo[q], (int)m);
        if (spaced.n) key = spaced_canonical(key, spaced);  // masked after selection
        keys[i] = key;
        key_win[i] = S;
      }
      at += (uint64_t)__popcll(b);
    }
  }
}

void launch_ocm_select(const uint64_t *codes, const uint64_t *invalid, const uint64_t *pos_off,
                       uint64_t n_reads, uint64_t positions, uint32_t k, uint32_t m,
                       uint64_t seed, const uint64_t *inner_table, uint8_t *flag,
                       uint8_t *delta, uint32_t *block_counts, uint64_t *totals,
                       hipStream_t s) {
  (void)pos_off;
  (void)n_reads;
  (void)totals;
  uint64_t blocks = (positions + kOcmTile - 1) / kOcmTile;
  if (!blocks) return;
  ocm_select_kernel<<<(uint32_t)blocks, 256, 0, s>>>(codes, invalid, positions, k, m, seed,
                                                     inner_table, flag, delta, block_counts);
}

void launch_ocm_compact(const uint64_t *codes, const uint8_t *flag, const uint8_t *delta,
                        const uint64_t *block_off, uint64_t positions, uint32_t m,
                        const SpacedRuns &spaced, uint64_t *keys, uint64_t *key_pos,
                        hipStream_t s) {
  const uint64_t tiles = (positions + kOcmTile - 1) / kOcmTile;
  if (!tiles) return;
  // a wave per tile, at most 4096 blocks of four waves
  const uint64_t blocks = std::min<uint64_t>((tiles + 3) / 4, 4096);
  ocm_compact_kernel<<<(uint32_t)blocks, 256, 0, s>>>(codes, flag, delta, block_off, positions,
                                                      m, spaced, keys, key_pos);
}

// ---------------------------------------------------------------------------
// Long k-mers (k > 32, scan_long_kmers, src/lib.rs:683-730). Forward hash
// sum (c_i + 1) B^(k-1-i), reverse-complement hash sum (4 - c_i) B^i, both
// mod 2^64; the key is splitmix64(min ^ domain). One thread per 64 window
// starts: the first valid window of a streak is hashed directly in O(k), the
// next ones roll in O(1).
// ---------------------------------------------------------------------------
constexpr uint64_t kKmerHashBase = 0x9e3779b185ebca87ULL;

LongKmerParams long_kmer_params(uint32_t k) {
  LongKmerParams lp{};
  uint64_t inv = kKmerHashBase;
  for (int i = 0; i < 6; i++) inv *= 2 - kKmerHashBase * inv;  // Newton steps
  uint64_t power = 1, b = kKmerHashBase;
  for (uint32_t e = k - 1; e; e >>= 1) {
    if (e & 1) power *= b;
    b *= b;
  }
  lp.power = power;
  lp.inverse = inv;
  lp.domain = splitmix64((uint64_t)k ^ 0x6b6d65722d763031ULL);
  lp.k = k;
  return lp;
}

__device__ __forceinline__ uint64_t code_at(const uint64_t *codes, uint64_t p) {
  return (codes[p >> 5] >> (2 * (p & 31))) & 3;
}

__global__ void long_kmer_keys_kernel(const uint64_t *codes, const uint64_t *invalid,
                                      uint64_t n_words, LongKmerParams lp, uint64_t seed,
                                      uint64_t sampling, uint64_t *keep, uint32_t *counts,
                                      uint64_t *pos_key) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint64_t valid = ~invalid_window_starts(invalid, w, lp.k);
  uint64_t kept = 0, fwd = 0, rev = 0;
  bool have = false;
  for (int i = 0; i < 64; i++) {
    if (!((valid >> i) & 1)) {
      have = false;
      continue;
    }
    const uint64_t p = w * 64 + i;
    if (!have) {
      fwd = 0;
      rev = 0;
      uint64_t weight = 1;
      for (uint32_t j = 0; j < lp.k; j++) {
        const uint64_t c = code_at(codes, p + j);
        fwd = fwd * kKmerHashBase + c + 1;
        rev += (4 - c) * weight;
        weight *= kKmerHashBase;
      }
      have = true;
    } else {
      const uint64_t out = code_at(codes, p - 1), in = code_at(codes, p + lp.k - 1);
      fwd = (fwd - (out + 1) * lp.power) * kKmerHashBase + in + 1;
      rev = (rev - (4 - out)) * lp.inverse + (4 - in) * lp.power;
    }
    const uint64_t key = splitmix64((fwd < rev ? fwd : rev) ^ lp.domain);
    if (sampling && mixsplit(key, seed) % sampling) continue;  // not a modimizer
    kept |= 1ULL << i;
    pos_key[p] = key;
  }
  keep[w] = kept;
  counts[w] = (uint32_t)__popcll(kept);
}

__global__ void compact_kept_kernel(const uint64_t *keep, const uint64_t *off,
                                    const uint64_t *pos_key, uint64_t n_words, uint64_t *keys,
                                    uint64_t *key_win) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint64_t kb = keep[w];
  uint64_t at = off[w];
  for (uint64_t m = kb; m; m &= m - 1) {
    const uint64_t p = w * 64 + __builtin_ctzll(m);
    keys[at] = pos_key[p];
    key_win[at] = p;
    at++;
  }
}

void launch_long_kmer_keys(const uint64_t *codes, const uint64_t *invalid, uint64_t positions,
                           const LongKmerParams &lp, uint64_t seed, uint64_t sampling,
                           uint64_t *keep, uint32_t *counts, uint64_t *pos_key, hipStream_t s) {
  const uint64_t words = (positions + 63) / 64;
  if (!words) return;
  long_kmer_keys_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(
      codes, invalid, words, lp, seed, sampling, keep, counts, pos_key);
}

void launch_compact_kept(const uint64_t *keep, const uint64_t *off, const uint64_t *pos_key,
                         uint64_t positions, uint64_t *keys, uint64_t *key_win, hipStream_t s) {
  const uint64_t words = (positions + 63) / 64;
  if (!words) return;
  compact_kept_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(keep, off, pos_key, words,
                                                                      keys, key_win);
}

// ---------------------------------------------------------------------------
// Randstrobemers (src/anchors.rs:136-351). Candidates are the valid k-mer
// windows, or the selected minimizer starts (each once); both query
// orientations are linked from the forward candidates, since the reverse
// complement holds the same canonical strobes in mirrored order.
// ---------------------------------------------------------------------------
__global__ void popc_words_kernel(const uint64_t *words, uint64_t n, uint32_t *counts) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (w < n) counts[w] = (uint32_t)__popcll(words[w]);
}

__global__ void strobe_kmer_candidates_kernel(const uint64_t *invalid, uint64_t n_words,
                                              uint32_t k, uint64_t *cand, uint32_t *counts) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint64_t v = ~invalid_window_starts(invalid, w, k);
  cand[w] = v;
  counts[w] = (uint32_t)__popcll(v);
}

__global__ void strobe_mark_selected_kernel(const uint8_t *flag, const uint8_t *delta,
                                            uint64_t positions, uint64_t *cand) {
  uint64_t S = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (S >= positions || !flag[S]) return;
  const uint64_t p = S + delta[S];
  atomicOr((unsigned long long *)&cand[p >> 6], (unsigned long long)(1ULL << (p & 63)));
}

// Candidate i in position order: start, strobe hash, run id (the number of
// invalid positions before it; equal ids = same run of valid bases).
__global__ void strobe_compact_kernel(const uint64_t *cand, const uint64_t *cand_off,
                                      const uint64_t *invalid, const uint64_t *inv_off,
                                      const uint64_t *codes, uint64_t n_words, uint32_t len,
                                      SpacedRuns spaced, uint64_t *cpos, uint64_t *chash,
                                      uint64_t *crun) {
  uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint64_t c = cand[w], inv = invalid[w];
  for (uint64_t m = c; m; m &= m - 1) {
    const int i = __builtin_ctzll(m);
    const uint64_t below = i ? (~0ULL >> (64 - i)) : 0;
    const uint64_t at = cand_off[w] + __popcll(c & below);
    const uint64_t p = w * 64 + i;
    uint64_t key = canonical_from_window(stream_bits(codes, 2 * p), (int)len);
    if (spaced.n) key = spaced_canonical(key, spaced);
    cpos[at] = p;
    chash[at] = strobe_hash(key);
    crun[at] = inv_off[w] + __popcll(inv & below);
  }
}

// keys[i]: the forward key of anchor i (partner: the later candidate of least
// rank, first on ties); keys[n + i]: its key in the reverse-complement
// orientation (partner: the earlier candidate of least rank, nearest on
// ties). Partners lie at a distance in [lo, hi] in the same run; an anchor
// without one pairs with itself.
__global__ void strobe_link_kernel(const uint64_t *cpos, const uint64_t *chash,
                                   const uint64_t *crun, uint64_t n, uint64_t lo, uint64_t hi,
                                   uint64_t *keys) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t p = cpos[i], h = chash[i], run = crun[i];
  uint64_t partner = h;
  uint32_t best = 64;
  for (uint64_t j = i + 1; j < n; j++) {
    const uint64_t d = cpos[j] - p;
    if (d > hi || crun[j] != run) break;
    if (d < lo) continue;
    const uint32_t r = strobe_rank(h, chash[j]);
    if (r < best) {
      best = r;
      partner = chash[j];
      if (!r) break;
    }
  }
  keys[i] = strobe_pair(h, partner);
  partner = h;
  best = 64;
  for (uint64_t j = i; j-- > 0;) {
    const uint64_t d = p - cpos[j];
    if (d > hi || crun[j] != run) break;
    if (d < lo) continue;
    const uint32_t r = strobe_rank(h, chash[j]);
    if (r < best) {
      best = r;
      partner = chash[j];
      if (!r) break;
    }
  }
  keys[n + i] = strobe_pair(h, partner);
}

// key_off[n + r] = total + key_off[r] for r in 1..n (key_off[n] == total).
__global__ void strobe_offsets_kernel(uint64_t *key_off, uint64_t n_reads, uint64_t total) {
  uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x + 1;
  if (r <= n_reads) key_off[n_reads + r] = total + key_off[r];
}

// counts[r][c] = max over the two orientations (rows r and n + r).
__global__ void strand_max_kernel(uint32_t *counts, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t a = counts[i], b = counts[n + i];
    if (b > a) counts[i] = b;
  }
}

void launch_popc_words(const uint64_t *words, uint64_t n, uint32_t *counts, hipStream_t s) {
  if (!n) return;
  popc_words_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(words, n, counts);
}
void launch_strobe_kmer_candidates(const uint64_t *invalid, uint64_t positions, uint32_t k,
                                   uint64_t *cand, uint32_t *counts, hipStream_t s) {
  const uint64_t words = (positions + 63) / 64;
  if (!words) return;
  strobe_kmer_candidates_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(invalid, words, k,
                                                                                  cand, counts);
}
void launch_strobe_mark_selected(const uint8_t *flag, const uint8_t *delta, uint64_t positions,
                                 uint64_t *cand, hipStream_t s) {
  if (!positions) return;
  strobe_mark_selected_kernel<<<(uint32_t)((positions + 255) / 256), 256, 0, s>>>(
      flag, delta, positions, cand);
}
void launch_strobe_compact(const uint64_t *cand, const uint64_t *cand_off, const uint64_t *invalid,
                           const uint64_t *inv_off, const uint64_t *codes, uint64_t positions,
                           uint32_t len, const SpacedRuns &spaced, uint64_t *cpos,
                           uint64_t *chash, uint64_t *crun, hipStream_t s) {
  const uint64_t words = (positions + 63) / 64;
  if (!words) return;
  strobe_compact_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(
      cand, cand_off, invalid, inv_off, codes, words, len, spaced, cpos, chash, crun);
}
void launch_strobe_link(const uint64_t *cpos, const uint64_t *chash, const uint64_t *crun,
                        uint64_t n, uint64_t lo, uint64_t hi, uint64_t *keys, hipStream_t s) {
  if (!n) return;
  strobe_link_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(cpos, chash, crun, n, lo, hi,
                                                                   keys);
}
void launch_strobe_offsets(uint64_t *key_off, uint64_t n_reads, uint64_t total, hipStream_t s) {
  if (!n_reads) return;
  strobe_offsets_kernel<<<(uint32_t)((n_reads + 255) / 256), 256, 0, s>>>(key_off, n_reads,
                                                                            total);
}
void launch_strand_max(uint32_t *counts, uint64_t n, hipStream_t s) {
  if (!n) return;
  strand_max_kernel<<<grid_for(n, 256, 16384), 256, 0, s>>>(counts, n);
}

// ---------------------------------------------------------------------------
// The lookup (storage.rs:2761-2925 on the plane layout)
// ---------------------------------------------------------------------------
template <int B>
struct FieldCfg {
  static constexpr int VB = B == 1 ? 2 : B == 2 ? 4 : B == 4 ? 8 : 16;  // bytes per lane
  static constexpr int F = VB * 8 / B;                                   // fields per lane
  static constexpr int NW = VB >= 4 ? VB / 4 : 1;                        // words per lane
  // 16-bit words for 2-byte lanes: the loaded value needs no widening, so no
  // wait is placed right after a load
  using Word = std::conditional_t<VB == 2, uint16_t, uint32_t>;
};

// VB bytes at p, which is VB-aligned: device rows start on 16-byte
// boundaries (the host pads the row pitch), lanes read VB-byte slices of them.
template <int B>
__device__ __forceinline__ void load_lane(const uint8_t *p, bool nontemporal,
                                          typename FieldCfg<B>::Word (&x)[FieldCfg<B>::NW]) {
  constexpr int VB = FieldCfg<B>::VB;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  if constexpr (VB == 16) {
    if (nontemporal) {
      // rows are read once per seed: stream them past the caches
      u32x4 v = __builtin_nontemporal_load((const u32x4 *)p);
      x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    } else {
      uint4 v = *(const uint4 *)p;
      x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    }
  } else if constexpr (VB == 8) {
    uint2 v = *(const uint2 *)p;
    x[0] = v.x; x[1] = v.y;
  } else if constexpr (VB == 4) {
    x[0] = *(const uint32_t *)p;
  } else {
    (void)nontemporal;
    x[0] = *(const uint16_t *)p;
  }
}

// Bit f of the result: field f of the lane's bytes equals fp.
template <int B>
__device__ __forceinline__ uint32_t field_hits(const typename FieldCfg<B>::Word (&x)[FieldCfg<B>::NW],
                                               uint32_t fp) {
  uint32_t m = 0;
  if constexpr (B == 8) {
#pragma unroll
    for (int w = 0; w < 4; w++) {
      uint32_t v = x[w] ^ (fp * 0x01010101u);
      uint32_t z = ~(((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
      // gather bits 7, 15, 23, 31 into bits 0..3: (z >> 7) * (1 + 2^7 + 2^14 + 2^21)
      // puts byte i's bit at 21 + i, with no carry below bit 25
      m |= (((z >> 7) * 0x00204081u) >> 21 & 0xFu) << (4 * w);
    }
  } else if constexpr (B == 16) {
#pragma unroll
    for (int w = 0; w < 4; w++) {
      m |= (uint32_t)((x[w] & 0xFFFFu) == fp) << (2 * w);
      m |= (uint32_t)((x[w] >> 16) == fp) << (2 * w + 1);
    }
  } else if constexpr (B == 32) {
#pragma unroll
    for (int w = 0; w < 4; w++) m |= (uint32_t)(x[w] == fp) << w;
  } else if constexpr (B == 4) {
#pragma unroll
    for (int w = 0; w < 2; w++) {
      uint32_t v = x[w] ^ (fp * 0x11111111u);
      uint32_t z = ~(((v & 0x77777777u) + 0x77777777u) | v) & 0x88888888u;
#pragma unroll
      for (int i = 0; i < 8; i++) m |= ((z >> (4 * i + 3)) & 1u) << (8 * w + i);
    }
  } else if constexpr (B == 2) {
    uint32_t v = x[0] ^ (fp * 0x55555555u);
    uint32_t z = ~(v | (v >> 1)) & 0x55555555u;
#pragma unroll
    for (int i = 0; i < 16; i++) m |= ((z >> (2 * i)) & 1u) << i;
  } else {
    m = (fp ? x[0] : ~x[0]) & 0xFFFFu;
  }
  return m;
}

template <int B, int T, int D>
__global__ __launch_bounds__(256) void lookup_kernel(LookupArgs a) {
  using C = FieldCfg<B>;
  using LaneWord = typename C::Word;
  constexpr int F = C::F;
  constexpr int NW = C::NW;
  constexpr int VB = C::VB;
  constexpr int HB = 4;  // rows loaded per block of hashes
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave_slot = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const BankDev &bk = a.bank;
  const uint32_t tile_fields = 64u * F;
  const uint32_t H = bk.hashes;
  const uint64_t kmask = a.k >= 32 ? ~0ULL : ((1ULL << (2 * a.k)) - 1);
  (void)kmask;
  const uint64_t win_mask = a.k >= 64 ? ~0ULL : ((1ULL << a.k) - 1);
  const uint64_t len_mask = bk.geo.length - 1;
  const uint64_t seg_stride = bk.geo.length * bk.row_bytes;

  // Field counters: per key, the F hit bits of a lane are added as byte
  // counters, four fields per word (bit i of a nibble spread to byte i by one
  // multiply), and spilled into the 32-bit counters every 255 keys and at
  // every flush.
  constexpr int Q = F / 4;
  uint32_t acc[T][F], acc8[T][Q];
#pragma unroll
  for (int t = 0; t < T; t++) {
#pragma unroll
    for (int f = 0; f < F; f++) acc[t][f] = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) acc8[t][q] = 0;
  }
  uint32_t pending = 0;  // keys in acc8 (wave-uniform, below 255)
  auto spill = [&]() {
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
      for (int q = 0; q < Q; q++) {
#pragma unroll
        for (int i = 0; i < 4; i++) acc[t][4 * q + i] += (acc8[t][q] >> (8 * i)) & 0xFFu;
        acc8[t][q] = 0;
      }
    pending = 0;
  };
  int64_t held_tg = -1;     // aggregate mode: tile group whose sums acc holds
  uint64_t held_keys = 0;

  uint32_t col0[T], fmask[T], mmask[T];
  // Recovered singleton colours of the current key, as field bits per lane
  // and tile (one wave's slice; written and read by that wave only).
  __shared__ uint32_t rec_bits_all[4 * T * 64];
  uint32_t *rec_bits = rec_bits_all + (threadIdx.x >> 6) * T * 64;

  auto flush_aggregate = [&]() {
    spill();
#pragma unroll
    for (int t = 0; t < T; t++) {
#pragma unroll
      for (int f = 0; f < F; f++) {
        if (acc[t][f]) {
          uint32_t color = bk.col_color[col0[t] + f];
          atomicAdd((unsigned long long *)&a.partial[wave_slot * a.n_colors + color],
                    (unsigned long long)acc[t][f]);
          acc[t][f] = 0;
        }
      }
    }
  };

  // Work slices: block x takes the units of slice x % slices, a contiguous
  // range, from that slice's own counter (128 B apart). Blocks are dealt to
  // the 8 XCDs in turn, so with 8 slices each XCD's L2 serves one range of a
  // sorted seed list, where neighbouring seeds share rows.
  const uint32_t n_slices = a.slices > 1 ? a.slices : 1;
  const uint32_t slice = blockIdx.x % n_slices;
  const uint32_t u_lo = (uint32_t)((uint64_t)a.n_units * slice / n_slices);
  const uint32_t u_hi = (uint32_t)((uint64_t)a.n_units * (slice + 1) / n_slices);
  while (true) {
    uint32_t unit = 0;
    if (lane == 0) unit = atomicAdd(a.queue + 32 * slice, 1u);
    unit = u_lo + readlane32(unit, 0);
    if (unit >= u_hi) break;
    const uint64_t sp = unit / a.n_tg;
    const uint32_t tg = unit % a.n_tg;
    const uint64_t i0 = a.begin + sp * (uint64_t)a.span;
    uint64_t i1 = i0 + a.span;
    if (i1 > a.end) i1 = a.end;

    if (!a.records && held_tg != (int64_t)tg) {
      if (held_tg >= 0) flush_aggregate();
      held_tg = tg;
      held_keys = 0;
    }
#pragma unroll
    for (int t = 0; t < T; t++) {
      uint32_t tile = tg * T + t;
      col0[t] = tile * tile_fields + lane * F;
      uint32_t fm = 0, mm = 0;
      if (tile < bk.n_tiles) {
#pragma unroll
        for (int f = 0; f < F; f++) {
          uint32_t col = col0[t] + f;
          if (col < bk.groups) {
            fm |= 1u << f;
            if (col >= bk.first_merged) mm |= 1u << f;
          }
        }
      }
      fmask[t] = fm;
      mmask[t] = mm;
    }

    uint64_t r = 0, seg_end = i1;
    if (a.records) {
      r = find_segment(a.seg_off, a.n_reads, i0);
      seg_end = a.seg_off[r + 1] < i1 ? a.seg_off[r + 1] : i1;
    }
    auto flush_read = [&]() {
      spill();
#pragma unroll
      for (int t = 0; t < T; t++) {
#pragma unroll
        for (int f = 0; f < F; f++) {
          if (acc[t][f]) {
            uint32_t color = bk.col_color[col0[t] + f];
            atomicAdd(&a.counts[r * a.n_colors + color], acc[t][f]);
            acc[t][f] = 0;
          }
        }
      }
    };
    auto fd_member = [&](uint64_t kidx, uint32_t slot) {
      const uint32_t bit = a.fd_member_bit + slot;
      atomicOr(a.fd_hits + kidx * a.fd_row_words + a.fd_off + (bit >> 5), 1u << (bit & 31));
    };
    auto out_add = [&](uint32_t color) {
      if (a.records) atomicAdd(&a.counts[r * a.n_colors + color], 1u);
      else atomicAdd((unsigned long long *)&a.partial[wave_slot * a.n_colors + color], 1ULL);
    };

    for (uint64_t c = i0; c < i1; c += 64) {
      const uint64_t idx = c + lane;
      bool v = idx < i1;
      uint64_t key = 0;
      if (a.source == kSeedKmers) {
        if (v) {
          uint64_t iv = stream_bits(a.invalid, idx);
          v = (iv & win_mask) == 0;
          key = canonical_from_window(stream_bits(a.codes, 2 * idx), (int)a.k);
          if (a.spaced.n) key = spaced_canonical(key, a.spaced);
        }
      } else if (v) {
        key = a.keys[idx];
      }
      if (v && bk.partitions > 1) v = partition_id(key, a.seed, bk.partitions) == bk.partition;
      uint32_t fp = v ? (uint32_t)fingerprint(key, a.seed, a.bits) : 0;
      // PositionHash::new for this lane's key, broadcast per key below
      const uint64_t lane_first = mixsplit(key, a.seed);
      const uint64_t lane_step = splitmix64(lane_first);
      // byte offset of the key's first segment; rows i follow at i segments
      const uint64_t lane_rowb = slot_base(lane_first, bk.geo) * bk.row_bytes;
      if (a.sampling && lane_first % a.sampling) v = false;  // not a modimizer
      uint32_t rec_start = 0, rec_len = 0;
      if (a.rec.enabled && v) {
        uint32_t h = abandoned_hash(key, a.seed);
        uint64_t cap_mask = (1ULL << a.rec.cap_log2) - 1;
        uint64_t slot = a.rec.cap_log2 ? (uint64_t)(h >> (32 - a.rec.cap_log2)) : 0;
        while (true) {
          uint64_t e = a.rec.dir[slot];
          uint32_t pi = (uint32_t)e;
          if (pi == kNone) break;
          if ((uint32_t)(e >> 32) == h) {
            rec_len = a.rec.post[pi];
            rec_start = pi + 1;
            break;
          }
          slot = (slot + 1) & cap_mask;
        }
      }
      uint64_t vmask = __ballot(v);
      held_keys += __popcll(vmask);

      while (vmask) {
        int jj[D];
        uint64_t kk[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
          if (vmask) {
            jj[d] = __builtin_ctzll(vmask);
            vmask &= vmask - 1;
            kk[d] = readlane64(key, jj[d]);
          } else {
            jj[d] = -1;
            kk[d] = 0;
          }
        }
        // Issue the row loads of all D keys before consuming any.
        LaneWord x[D][T][NW];
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int t = 0; t < T; t++)
#pragma unroll
            for (int w = 0; w < NW; w++) x[d][t][w] = 0;
        // PositionHash per key, advanced one row at a time: h_i = first + i*step,
        // row i at (base + i*len + ((h_i ^ h_i >> 33) & (len - 1))) * row_bytes
        uint64_t hv[D], step[D], rowb[D];
        uint32_t rl[D], rs[D], pre[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
          int l = jj[d] < 0 ? 0 : jj[d];
          hv[d] = readlane64(lane_first, l);
          step[d] = readlane64(lane_step, l);
          rowb[d] = readlane64(lane_rowb, l);
          // the first 64 recovery postings of the key travel with its rows
          rl[d] = jj[d] < 0 ? 0 : readlane32(rec_len, l);
          rs[d] = readlane32(rec_start, l);
          pre[d] = lane < rl[d] ? a.rec.post_col[rs[d] + lane] : kNone;
        }
        for (uint32_t h0 = 0; h0 < H; h0 += HB) {
          LaneWord raw[D][HB][T][NW];
          // The load condition is wave-uniform: lanes past the last column
          // read row padding (a bank buffer ends with 4 KiB of zeros), and
          // their fields are masked out below. Non-temporal or plain loads are
          // chosen once per block of rows.
          auto issue = [&](auto nt_tag) {
            constexpr bool NT = decltype(nt_tag)::value;
#pragma unroll
            for (int d = 0; d < D; d++) {
#pragma unroll
              for (int hh = 0; hh < HB; hh++) {
                const bool live = jj[d] >= 0 && h0 + hh < H;
                const uint8_t *row =
                    bk.data + rowb[d] + ((hv[d] ^ (hv[d] >> 33)) & len_mask) * bk.row_bytes;
#pragma unroll
                for (int t = 0; t < T; t++) {
                  const uint32_t tile = tg * T + t;
                  if (live && tile < bk.n_tiles) {
                    load_lane<B>(row + (uint64_t)tile * 64 * VB + lane * VB, NT, raw[d][hh][t]);
                  } else {
#pragma unroll
                    for (int w = 0; w < NW; w++) raw[d][hh][t][w] = 0;
                  }
                }
                hv[d] += step[d];
                rowb[d] += seg_stride;
              }
            }
          };
          if (VB == 16 && a.nontemporal) issue(std::true_type{});
          else issue(std::false_type{});
#pragma unroll
          for (int d = 0; d < D; d++)
#pragma unroll
            for (int hh = 0; hh < HB; hh++)
#pragma unroll
              for (int t = 0; t < T; t++)
#pragma unroll
                for (int w = 0; w < NW; w++) x[d][t][w] ^= raw[d][hh][t][w];
        }
#pragma unroll
        for (int d = 0; d < D; d++) {
          if (jj[d] < 0) continue;
          const uint64_t kidx = c + (uint64_t)jj[d];
          if (a.records) {
            while (kidx >= seg_end) {
              flush_read();
              r++;
              seg_end = a.seg_off[r + 1] < i1 ? a.seg_off[r + 1] : i1;
            }
          }
          const uint32_t fpd = readlane32(fp, jj[d]);
          uint32_t m[T];
#pragma unroll
          for (int t = 0; t < T; t++) m[t] = field_hits<B>(x[d][t], fpd) & fmask[t];
          // Recovery: OR recovered colours into this key's hits
          // (storage.rs:2771-2787, 2916-2924). The key's postings are spread
          // over the lanes; singleton columns set their field bit through the
          // wave's LDS slice, merged members are settled by the posting's lane.
          if (rl[d]) {
#pragma unroll
            for (int t = 0; t < T; t++) rec_bits[t * 64 + lane] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (uint32_t q0 = 0; q0 < rl[d]; q0 += 64) {
              const uint32_t q = q0 + lane;
              const uint32_t e =
                  q0 == 0 ? pre[d] : (q < rl[d] ? a.rec.post_col[rs[d] + q] : kNone);
              const uint32_t col = e & 0xFFFFFFu;
              const uint32_t rel = col / tile_fields - tg * T;  // tile within the group
              const bool mine = e != kNone && (e >> 24) == bk.colmap && rel < (uint32_t)T;
              const uint32_t within = col % tile_fields;
              if (mine && col < bk.first_merged)
                atomicOr(&rec_bits[rel * 64 + within / F], 1u << (within % F));
              if (__ballot(mine && col >= bk.first_merged)) {
                // the group's fingerprint hit, from the lane that owns the field
                uint32_t mo = 0;
#pragma unroll
                for (int t = 0; t < T; t++) {
                  uint32_t v = __shfl(m[t], (int)(within / F));
                  if (rel == (uint32_t)t) mo = v;
                }
                if (mine && col >= bk.first_merged) {
                  const uint32_t o = a.rec.post[rs[d] + q];
                  bool hit = false;
                  if ((mo >> (within % F)) & 1u) {
                    uint64_t bit = bk.mg_mask_bit[col - bk.first_merged] + bk.color_bit[o];
                    hit = xor_rows_bits(bk, kk[d], a.seed, bit, 1) & 1;
                  }
                  if (!hit) {
                    if (a.fd_hits)
                      fd_member(kidx, bk.mg_off[col - bk.first_merged] + bk.color_bit[o]);
                    else
                      out_add(o);
                  }
                }
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < T; t++) m[t] |= rec_bits[t * 64 + lane];
          }
          // Multi-member groups: decode the XORed membership mask.
#pragma unroll
          for (int t = 0; t < T; t++) {
            uint32_t mh = m[t] & mmask[t];
            m[t] &= ~mmask[t];
            while (mh) {
              int f = __builtin_ctz(mh);
              mh &= mh - 1;
              uint32_t g = col0[t] + f - bk.first_merged;
              uint32_t lo = bk.mg_off[g], width = bk.mg_off[g + 1] - lo;
              uint64_t mask = xor_rows_bits(bk, kk[d], a.seed, bk.mg_mask_bit[g], width);
              while (mask) {
                int b = __builtin_ctzll(mask);
                mask &= mask - 1;
                if (a.fd_hits) fd_member(kidx, lo + b);
                else out_add(bk.mg_members[lo + b]);
              }
            }
          }
          if (a.fd_hits) {
            // this seed's singleton hits: F bits per lane at (tile * 64 + lane) * F
            uint8_t *row = (uint8_t *)(a.fd_hits + kidx * a.fd_row_words + a.fd_off);
#pragma unroll
            for (int t = 0; t < T; t++) {
              const uint32_t tile = tg * T + t;
              if constexpr (F == 4) {
                const uint32_t up = __shfl_down(m[t], 1);
                if (tile < bk.n_tiles && !(lane & 1))
                  row[(tile * 64 + lane) / 2] = (uint8_t)((m[t] & 0xFu) | (up & 0xFu) << 4);
              } else if (tile < bk.n_tiles) {
                if constexpr (F == 8) row[tile * 64 + lane] = (uint8_t)m[t];
                else ((uint16_t *)row)[tile * 64 + lane] = (uint16_t)m[t];
              }
            }
          } else {
#pragma unroll
            for (int t = 0; t < T; t++)
#pragma unroll
              for (int q = 0; q < Q; q++)
                acc8[t][q] += (((m[t] >> (4 * q)) & 0xFu) * 0x00204081u) & 0x01010101u;
            if (++pending == 255) spill();
          }
        }
      }
    }
    if (a.records) {
      flush_read();
    } else if (held_keys > (1ULL << 30)) {
      flush_aggregate();
      held_keys = 0;
    }
  }
  if (!a.records && held_tg >= 0) flush_aggregate();
}

template <int B, int T, int D>
static void launch_one(const LookupArgs &a, uint32_t blocks, hipStream_t s) {
  lookup_kernel<B, T, D><<<blocks, 256, 0, s>>>(a);
}

// Deeper prefetch (8, 16 seeds in flight) only for narrow lanes (b <= 4):
// small rows need more loads in flight per wave to cover HBM latency.
template <int B, int T>
static void launch_d(const LookupArgs &a, const KernelConfig &c, hipStream_t s) {
  if constexpr (B <= 4) {
    if (c.prefetch == 8) return launch_one<B, T, 8>(a, c.blocks, s);
    if (c.prefetch == 16) return launch_one<B, T, 16>(a, c.blocks, s);
  }
  switch (c.prefetch) {
    case 1: launch_one<B, T, 1>(a, c.blocks, s); break;
    case 4: launch_one<B, T, 4>(a, c.blocks, s); break;
    default: launch_one<B, T, 2>(a, c.blocks, s); break;
  }
}

template <int B>
static void launch_t(const LookupArgs &a, const KernelConfig &c, hipStream_t s) {
  if (c.tiles_per_wave == 2) launch_d<B, 2>(a, c, s);
  else launch_d<B, 1>(a, c, s);
}

void launch_lookup(const LookupArgs &a, const KernelConfig &c, hipStream_t s) {
  switch (c.bits) {
    case 1: launch_t<1>(a, c, s); break;
    case 2: launch_t<2>(a, c, s); break;
    case 4: launch_t<4>(a, c, s); break;
    case 8: launch_t<8>(a, c, s); break;
    case 16: launch_t<16>(a, c, s); break;
    default: launch_t<32>(a, c, s); break;
  }
}

template <int B, int T, int D>
static int occupancy_one() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lookup_kernel<B, T, D>, 256, 0) !=
      hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

template <int B, int T>
static int occupancy_d(uint32_t d) {
  if constexpr (B <= 4) {
    if (d == 8) return occupancy_one<B, T, 8>();
    if (d == 16) return occupancy_one<B, T, 16>();
  }
  return d == 1 ? occupancy_one<B, T, 1>() : d == 4 ? occupancy_one<B, T, 4>()
                                                    : occupancy_one<B, T, 2>();
}

template <int B>
static int occupancy_t(const KernelConfig &c) {
  return c.tiles_per_wave == 2 ? occupancy_d<B, 2>(c.prefetch) : occupancy_d<B, 1>(c.prefetch);
}

int lookup_blocks_per_cu(const KernelConfig &c) {
  switch (c.bits) {
    case 1: return occupancy_t<1>(c);
    case 2: return occupancy_t<2>(c);
    case 4: return occupancy_t<4>(c);
    case 8: return occupancy_t<8>(c);
    case 16: return occupancy_t<16>(c);
    default: return occupancy_t<32>(c);
  }
}

// Findere reconstruction (findere::Scorer::sequence, src/findere.rs:74-104).
// One lane per region bit (colour), one wave per (64 bits, span of
// positions). Rows of invalid or unlooked-up s-mer windows are zero, so runs
// break at ambiguous bases and read separators by themselves; a wave starts
// z positions before its span to rebuild the run lengths it inherits.
constexpr uint32_t kFindereSpan = 2048;

__global__ __launch_bounds__(256) void findere_scan_kernel(
    const uint32_t *hits, uint64_t row_words, uint32_t region_off, uint32_t region_bits,
    const uint32_t *bit_color, const uint64_t *pos_off, uint64_t n_reads, uint64_t positions,
    uint32_t z, uint32_t records, uint32_t *counts, uint64_t *partial, uint32_t n_colors,
    uint64_t p_base) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint64_t groups = (region_bits + 63) / 64;
  const uint64_t spans = (positions - p_base + kFindereSpan - 1) / kFindereSpan;
  if (wave >= groups * spans) return;
  const uint64_t g = wave % groups, sp = wave / groups;
  const uint32_t j = (uint32_t)(g * 64 + lane);
  const uint32_t color = j < region_bits ? bit_color[j] : kNone;
  // positions [p_base, positions); hit rows are indexed by position - p_base
  const uint64_t p0 = p_base + sp * kFindereSpan;
  const uint64_t p1 = p0 + kFindereSpan < positions ? p0 + kFindereSpan : positions;
  const uint64_t start = p0 > p_base + z ? p0 - z : p_base;
  const uint32_t *col = hits + region_off + (j >> 5);
  const uint32_t sh = j & 31;
  const uint32_t window = z + 1;
  uint64_t r = find_segment(pos_off, n_reads, p0);
  uint64_t next = pos_off[r + 1];
  uint32_t run = 0, cnt = 0;
  for (uint64_t p = start; p < p1; p++) {
    const uint32_t x = color != kNone ? (col[(p - p_base) * row_words] >> sh) & 1u : 0u;
    run = x ? (run + 1 < window ? run + 1 : window) : 0;
    if (p < p0) continue;
    if (records && p >= next) {
      if (cnt) atomicAdd(&counts[r * n_colors + color], cnt);
      cnt = 0;
      while (pos_off[r + 1] <= p) r++;
      next = pos_off[r + 1];
    }
    cnt += run == window;
  }
  if (cnt) {
    if (records) atomicAdd(&counts[r * n_colors + color], cnt);
    else atomicAdd((unsigned long long *)&partial[color], (unsigned long long)cnt);
  }
}

void launch_findere_scan(const uint32_t *hits, uint64_t row_words, uint32_t region_off,
                         uint32_t region_bits, const uint32_t *bit_color,
                         const uint64_t *pos_off, uint64_t n_reads, uint64_t positions,
                         uint32_t z, uint32_t records, uint32_t *counts, uint64_t *partial,
                         uint32_t n_colors, hipStream_t s, uint64_t p_base) {
  if (positions <= p_base || !n_reads || !region_bits) return;
  const uint64_t waves =
      ((region_bits + 63) / 64) * ((positions - p_base + kFindereSpan - 1) / kFindereSpan);
  findere_scan_kernel<<<(uint32_t)((waves + 3) / 4), 256, 0, s>>>(
      hits, row_words, region_off, region_bits, bit_color, pos_off, n_reads, positions, z,
      records, counts, partial, n_colors, p_base);
}

// Random-row gather probe: whole rows of ROW = CH * 64 * LB bytes at
// uniformly random row indices, RPW rows per wave in flight, with no other
// work. Its rate is the ceiling the lookup's row gathers are measured against.
template <int LB, int CH, int RPW>
__global__ __launch_bounds__(256) void gather_probe_kernel(const uint8_t *data, uint64_t rows,
                                                           uint64_t n, uint64_t seed,
                                                           uint32_t *sink) {
  constexpr uint32_t ROW = CH * 64 * LB;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t acc = 0;
  for (uint64_t i0 = wave * RPW; i0 < n; i0 += waves * RPW) {
    uint32_t v[RPW][CH][LB / 4];
#pragma unroll
    for (int r = 0; r < RPW; r++) {
      const uint64_t i = i0 + r;
      const uint64_t row = mulhi64(splitmix64(i ^ seed), rows);
      const uint8_t *p = data + row * ROW + lane * LB;
#pragma unroll
      for (int c = 0; c < CH; c++) {
        if (i < n) {
          if constexpr (LB == 16) {
            u32x4 x = __builtin_nontemporal_load((const u32x4 *)(p + c * 64 * LB));
            v[r][c][0] = x.x; v[r][c][1] = x.y; v[r][c][2] = x.z; v[r][c][3] = x.w;
          } else {
            u32x2 x = __builtin_nontemporal_load((const u32x2 *)(p + c * 64 * LB));
            v[r][c][0] = x.x; v[r][c][1] = x.y;
          }
        } else {
#pragma unroll
          for (int w = 0; w < LB / 4; w++) v[r][c][w] = 0;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int c = 0; c < CH; c++)
#pragma unroll
        for (int w = 0; w < LB / 4; w++) acc ^= v[r][c][w];
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads live
}

// The rows of random keys under the bank's PositionHash geometry (H rows in H
// consecutive segments per key, storage.rs:3452-3481), KPW keys per wave in
// flight, no other work: the ceiling of the lookup's own row pattern.
template <int LB, int CH, int KPW>
__global__ __launch_bounds__(256) void gather_keys_kernel(const uint8_t *data, uint64_t row_bytes,
                                                          SegmentGeometry geo, uint32_t H,
                                                          uint64_t n_keys, uint64_t seed,
                                                          uint32_t *sink) {
  constexpr int HB = 8;
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t acc = 0;
  for (uint64_t i0 = wave * KPW; i0 < n_keys; i0 += waves * KPW) {
    for (uint32_t h0 = 0; h0 < H; h0 += HB) {
      uint32_t v[KPW][HB][CH][LB / 4];
#pragma unroll
      for (int q = 0; q < KPW; q++) {
        const uint64_t first = splitmix64((i0 + q) ^ seed), step = splitmix64(first);
        const uint64_t base = slot_base(first, geo);
#pragma unroll
        for (int hh = 0; hh < HB; hh++) {
          const uint32_t h = h0 + hh;
          const bool live = i0 + q < n_keys && h < H;
          const uint64_t slot = slot_at(base, first + (uint64_t)h * step, h, geo);
          const uint8_t *p = data + slot * row_bytes + lane * LB;
#pragma unroll
          for (int c = 0; c < CH; c++) {
            if (live) {
              if constexpr (LB == 16) {
                u32x4 x = __builtin_nontemporal_load((const u32x4 *)(p + c * 64 * LB));
                v[q][hh][c][0] = x.x; v[q][hh][c][1] = x.y; v[q][hh][c][2] = x.z; v[q][hh][c][3] = x.w;
              } else {
                u32x2 x = __builtin_nontemporal_load((const u32x2 *)(p + c * 64 * LB));
                v[q][hh][c][0] = x.x; v[q][hh][c][1] = x.y;
              }
            } else {
#pragma unroll
              for (int w = 0; w < LB / 4; w++) v[q][hh][c][w] = 0;
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < KPW; q++)
#pragma unroll
        for (int hh = 0; hh < HB; hh++)
#pragma unroll
          for (int c = 0; c < CH; c++)
#pragma unroll
            for (int w = 0; w < LB / 4; w++) acc ^= v[q][hh][c][w];
    }
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads live
}

bool launch_gather_keys(const uint8_t *data, uint64_t row_bytes, SegmentGeometry geo, uint32_t H,
                        uint64_t n_keys, uint32_t blocks, uint32_t *sink, hipStream_t s) {
  const uint64_t seed = 0x2545F4914F6CDD1DULL;
  switch (row_bytes) {
    case 512: gather_keys_kernel<8, 1, 1><<<blocks, 256, 0, s>>>(data, row_bytes, geo, H, n_keys, seed, sink); return true;
    case 1024: gather_keys_kernel<16, 1, 1><<<blocks, 256, 0, s>>>(data, row_bytes, geo, H, n_keys, seed, sink); return true;
    case 2048: gather_keys_kernel<16, 2, 1><<<blocks, 256, 0, s>>>(data, row_bytes, geo, H, n_keys, seed, sink); return true;
    default: return false;
  }
}

template <int LB, int CH>
static void launch_probe(const uint8_t *data, uint64_t rows, uint64_t n, uint32_t blocks,
                         uint32_t *sink, hipStream_t s) {
  gather_probe_kernel<LB, CH, 4><<<blocks, 256, 0, s>>>(data, rows, n, 0x5DEECE66DULL, sink);
}

bool launch_gather_probe(const uint8_t *data, uint64_t bytes, uint32_t row_bytes, uint64_t n,
                         uint32_t blocks, uint32_t *sink, hipStream_t s) {
  const uint64_t rows = bytes / row_bytes;
  if (!rows) return false;
  switch (row_bytes) {
    case 512: launch_probe<8, 1>(data, rows, n, blocks, sink, s); return true;
    case 1024: launch_probe<16, 1>(data, rows, n, blocks, sink, s); return true;
    case 2048: launch_probe<16, 2>(data, rows, n, blocks, sink, s); return true;
    case 4096: launch_probe<16, 4>(data, rows, n, blocks, sink, s); return true;
    case 8192: launch_probe<16, 8>(data, rows, n, blocks, sink, s); return true;
    default: return false;
  }
}

__global__ void reduce_partials_kernel(const uint64_t *partial, uint64_t waves,
                                       uint32_t n_colors, uint64_t *agg) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_colors) return;
  uint64_t s = 0;
  for (uint64_t w = 0; w < waves; w++) s += partial[w * n_colors + c];
  agg[c] += s;
}

void launch_reduce_partials(const uint64_t *partial, uint64_t waves, uint32_t n_colors,
                            uint64_t *agg, hipStream_t s) {
  if (!n_colors) return;
  reduce_partials_kernel<<<(n_colors + 255) / 256, 256, 0, s>>>(partial, waves, n_colors, agg);
}

}  // namespace zg

namespace zg {

__global__ void segment_bounds_kernel(const uint64_t *key_win, uint64_t n_keys,
                                      const uint64_t *pos_off, uint64_t n_reads,
                                      uint64_t *key_off) {
  uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (r > n_reads) return;
  uint64_t x = pos_off[r], lo = 0, hi = n_keys;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (key_win[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  key_off[r] = lo;
}

void launch_segment_bounds(const uint64_t *key_win, uint64_t n_keys, const uint64_t *pos_off,
                           uint64_t n_reads, uint64_t *key_off, hipStream_t s) {
  segment_bounds_kernel<<<(uint32_t)((n_reads + 1 + 255) / 256), 256, 0, s>>>(
      key_win, n_keys, pos_off, n_reads, key_off);
}

__global__ void diff_kernel(const uint64_t *off, uint64_t n, uint64_t *totals) {
  uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (r < n) totals[r] = off[r + 1] - off[r];
}

void launch_diff(const uint64_t *off, uint64_t n, uint64_t *totals, hipStream_t s) {
  if (!n) return;
  diff_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(off, n, totals);
}

// Column sums over a [n_reads x n_colors] matrix: blockIdx.y picks a slab of
// `slab` reads, threads of a row of blocks cover the colours (coalesced), one
// atomic per (slab, colour).
__global__ void colsum_kernel(const uint32_t *counts, uint64_t n_reads, uint32_t n_colors,
                              uint64_t slab, uint64_t *agg) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_colors) return;
  uint64_t r0 = (uint64_t)blockIdx.y * slab;
  uint64_t r1 = r0 + slab < n_reads ? r0 + slab : n_reads;
  uint64_t s = 0;
  for (uint64_t r = r0; r < r1; r++) s += counts[r * n_colors + c];
  if (s) atomicAdd((unsigned long long *)&agg[c], (unsigned long long)s);
}

void launch_colsum(const uint32_t *counts, uint64_t n_reads, uint32_t n_colors, uint64_t *agg,
                   hipStream_t s) {
  if (!n_colors || !n_reads) return;
  uint64_t slab = 512;
  if ((n_reads + slab - 1) / slab > 65535) slab = (n_reads + 65534) / 65535;
  uint64_t slabs = (n_reads + slab - 1) / slab;
  dim3 grid((n_colors + 255) / 256, (uint32_t)slabs);
  colsum_kernel<<<grid, 256, 0, s>>>(counts, n_reads, n_colors, slab, agg);
}

}  // namespace zg
