// zg_sweep.hip -- the batched query in bank order (the sorted sweep) on gfx950.
//
// A k-mer's H rows lie in H consecutive segments of L rows starting at its
// first segment mulhi(h, count), h = mixsplit(key, seed) (PositionHash,
// storage.rs:3452-3481). Per chunk of whole reads:
//   1. the chunk's valid windows become entries (h, position), put in first-
//      segment order by a stable LSD radix sort on the segment index (3
//      passes of 6 bits for 2^17 segments; tile histograms, a scan, and a
//      stable scatter per pass: no inter-block spinning);
//   2. bucket starts per segment;
//   3. the staged lookup: a block takes one segment and one 512-byte column
//      slice of the fingerprint plane, copies the H segments x L rows of that
//      slice into LDS (128 KiB at H=4, L=64), and looks up every entry of the
//      segment from LDS (XOR of H rows, b-bit compare), writing the hit bits
//      of its slice into the entry's hit row, indexed by position;
//   4. the per-read sum walks the hit rows in position order (sequential
//      reads) with bit-sliced counters and adds them to the read's counts.
// Each row is read from HBM about once per chunk instead of once per seed,
// the hit rows are written and read once, and nothing waits on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "zg_device.h"
#include "zg_internal.h"

namespace zg {

namespace {

constexpr uint32_t kSortThreads = 512;               // 8 waves
constexpr uint32_t kSortWaves = kSortThreads / 64;
constexpr uint32_t kSortRounds = 8;                  // rounds of 64 entries per wave
constexpr uint64_t kSortWaveItems = 64 * kSortRounds;
constexpr uint64_t kSortTile = kSortWaves * kSortWaveItems;  // 4096 entries per tile (48 KiB in LDS)
// Digit bits per radix pass (64 buckets: a tile's digit run averages 64
// entries, whole cache lines of the output)
constexpr uint32_t kMaxDigitBits = 6;
constexpr uint32_t kMaxBins = 1u << kMaxDigitBits;
constexpr uint32_t kStageThreads = 1024;             // 16 waves, one block per CU
constexpr uint32_t kStageWaves = kStageThreads / 64;
constexpr uint32_t kStageWindow = 128 * 1024;        // LDS rows of one segment window
constexpr uint32_t kSliceBytes = 512;                // plane bytes per slice (8 per lane)
constexpr uint32_t kStageOut = 512;                  // staged hit bytes per wave
constexpr uint64_t kSweepHitBytes = 160ULL << 30;    // hit rows of one chunk, at most
constexpr uint32_t kSumSpan = 4096;                  // positions per wave of the sum
constexpr uint32_t kXcds = 8;
// The bit-plane lookup at b >= 4 and b < 32: one 1024-thread block per CU
// with a 128 KiB window (H x L rows of one slice: 512 bytes for H <= 4,
// 256 for H = 8), so each hit row is written in halves; at b <= 2 and
// b = 32 (more registers per lane): two 512-thread blocks per CU with 64 KiB
// windows, one block's window fill overlapping the other's lookups.
constexpr bool plane_wide(int b) { return b >= 4 && b < 32; }
constexpr uint32_t plane_threads(int b) { return plane_wide(b) ? 1024 : 512; }
constexpr uint32_t plane_window(int b) { return plane_wide(b) ? 128 * 1024 : 64 * 1024; }

// The seeds of positions [p0, p1): canonical k-mers (k <= 32), optionally
// through a spaced seed, keyed by their first PositionHash word; modimizers
// keep the sampled ones (scan_modimizers, src/lib.rs:796-806).
struct PosSource {
  const uint64_t *codes, *invalid;
  uint64_t p0, n;  // chunk start, positions
  uint32_t k;
  uint64_t seed, sampling;
  SpacedRuns spaced;
  __device__ bool get(uint64_t i, uint64_t &h, uint32_t &pos) const {
    if (i >= n) return false;
    const uint64_t p = p0 + i;
    const uint64_t win = k >= 64 ? ~0ULL : (1ULL << k) - 1;
    if (stream_bits(invalid, p) & win) return false;
    uint64_t key = canonical_from_window(stream_bits(codes, 2 * p), (int)k);
    if (spaced.n) key = spaced_canonical(key, spaced);
    h = mixsplit(key, seed);
    if (sampling && h % sampling) return false;
    pos = (uint32_t)i;
    return true;
  }
};

// Explicit seed keys (minimizers, long k-mers, randstrobes) k0 .. k0 + n - 1.
struct KeySource {
  const uint64_t *keys;
  uint64_t k0, n;
  uint64_t seed;
  __device__ bool get(uint64_t i, uint64_t &h, uint32_t &pos) const {
    if (i >= n) return false;
    h = mixsplit(keys[k0 + i], seed);
    pos = (uint32_t)i;
    return true;
  }
};

// Entries of an earlier pass; the count is read from device memory.
struct EntSource {
  const uint64_t *h;
  const uint32_t *p;
  const uint64_t *n_ptr;
  uint64_t n;
  __device__ bool get(uint64_t i, uint64_t &hv, uint32_t &pos) const {
    if (i >= n) return false;
    hv = h[i];
    pos = p[i];
    return true;
  }
};

__device__ __forceinline__ void resolve(PosSource &) {}
__device__ __forceinline__ void resolve(KeySource &) {}
__device__ __forceinline__ void resolve(EntSource &s) { s.n = *s.n_ptr; }

// Tile t, wave w, round r, lane l holds item t*4096 + w*512 + r*64 + l:
// the tile order is the input order, which keeps the sort stable.
__device__ __forceinline__ uint64_t sort_item(uint64_t tile, uint32_t wave, uint32_t r,
                                              uint32_t lane) {
  return tile * kSortTile + wave * kSortWaveItems + (uint64_t)r * 64 + lane;
}

// The tile of a block: blocks go to the XCDs in turn, and each XCD takes a
// contiguous range of tiles, so the digit runs of neighbouring tiles (which
// are neighbours in the output) are written through the same L2.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t n) {
  const uint64_t x = b % kXcds, per = n / kXcds, rem = n % kXcds;
  return x * per + (x < rem ? x : rem) + b / kXcds;
}

template <class Src>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(Src src, uint64_t count, uint32_t shift,
                                                         uint32_t dbits, uint64_t n_tiles,
                                                         uint32_t *hist, uint64_t *listed) {
  __shared__ uint32_t bins[kMaxBins];
  resolve(src);
  const uint32_t nb = 1u << dbits, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < nb; i += kSortThreads) bins[i] = 0;
  __syncthreads();
  const uint64_t tile = xcd_tile(blockIdx.x, n_tiles);
  for (uint32_t r = 0; r < kSortRounds; r++) {
    const uint64_t i = sort_item(tile, wave, r, lane);
    uint64_t h = 0;
    uint32_t p = 0;
    const bool v = src.get(i, h, p);
    if (listed) {  // valid windows of the round's 64 positions
      const uint64_t m = __ballot(v);
      if (lane == 0 && i < src.n) listed[i >> 6] = m;
    }
    if (v) atomicAdd(&bins[(mulhi64(h, count) >> shift) & (nb - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < nb; d += kSortThreads) hist[(uint64_t)d * n_tiles + tile] = bins[d];
  if (tile == 0 && threadIdx.x == 0) hist[(uint64_t)nb * n_tiles] = 0;  // closes the scan input
}

// Stable scatter of one tile. Its items are ranked in LDS first: each
// wave's items follow the earlier waves' items of the same digit, in input
// order. The tile is then written out digit run by digit run, so that
// neighbouring threads write neighbouring entries (whole cache lines) after
// the same digit's entries of the earlier tiles.
// With `bounds` (the last pass), each entry that starts its segment in the
// output lowers bounds[segment] to its index: an entry starts it when it
// opens its tile's run of its digit (its predecessor is in another tile) or
// its predecessor in the run has another segment. bounds_fill_kernel then
// gives empty segments the start of the next one.
template <class Src>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(Src src, uint64_t count, uint32_t shift,
                                                            uint32_t dbits, uint64_t n_tiles,
                                                            const uint64_t *off, uint64_t *out_h,
                                                            uint32_t *out_p, uint32_t *bounds) {
  __shared__ uint32_t wbase[kSortWaves][kMaxBins];  // per wave: digit counts, then next LDS slot
  __shared__ uint32_t dstart[kMaxBins + 1];  // first LDS slot of each digit
  __shared__ uint64_t gbase[kMaxBins];       // global index of LDS slot 0 of each digit
  __shared__ uint64_t th[kSortTile];
  __shared__ uint32_t tp[kSortTile];
  resolve(src);
  const uint32_t nb = 1u << dbits, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = lane; i < nb; i += 64) wbase[wave][i] = 0;
  __syncthreads();
  const uint64_t tile = xcd_tile(blockIdx.x, n_tiles);
  uint64_t hv[kSortRounds];
  uint32_t pv[kSortRounds], dv[kSortRounds];
#pragma unroll
  for (uint32_t r = 0; r < kSortRounds; r++) {
    uint64_t h = 0;
    uint32_t p = 0;
    const bool v = src.get(sort_item(tile, wave, r, lane), h, p);
    hv[r] = h;
    pv[r] = p;
    dv[r] = v ? (uint32_t)((mulhi64(h, count) >> shift) & (nb - 1)) : kNone;
    if (v) atomicAdd(&wbase[wave][dv[r]], 1u);
  }
  __syncthreads();
  if (wave == 0) {  // nb <= 64: one lane per digit
    const uint32_t d = lane;
    uint32_t tot = 0;
    if (d < nb)
      for (uint32_t w = 0; w < kSortWaves; w++) tot += wbase[w][d];
    const uint64_t excl = wave_inclusive_scan(tot, lane) - tot;
    if (d < nb) {
      dstart[d] = (uint32_t)excl;
      gbase[d] = off[(uint64_t)d * n_tiles + tile] - excl;
      uint32_t run = (uint32_t)excl;
      for (uint32_t w = 0; w < kSortWaves; w++) {
        const uint32_t c = wbase[w][d];
        wbase[w][d] = run;
        run += c;
      }
      if (d == nb - 1) dstart[nb] = run;
    }
  }
  __syncthreads();
  const uint64_t below = lanemask_lt(lane);
#pragma unroll
  for (uint32_t r = 0; r < kSortRounds; r++) {
    const uint32_t d = dv[r];
    const bool v = d != kNone;
    // lanes holding the same digit: one ballot per digit bit
    uint64_t same = __ballot(v);
    for (uint32_t bit = 0; bit < dbits; bit++) {
      const uint64_t ones = __ballot((d >> bit) & 1);
      same &= ((d >> bit) & 1) ? ones : ~ones;
    }
    if (v) {
      const uint32_t at = wbase[wave][d] + (uint32_t)__popcll(same & below);
      th[at] = hv[r];
      tp[at] = pv[r];
      if (!(same & below)) wbase[wave][d] += (uint32_t)__popcll(same);  // the group's first lane
    }
  }
  __syncthreads();
  const uint32_t n_tile = dstart[nb];
  for (uint32_t i = threadIdx.x; i < n_tile; i += kSortThreads) {
    const uint64_t h = th[i];
    const uint64_t seg = mulhi64(h, count);
    const uint32_t d = (uint32_t)((seg >> shift) & (nb - 1));
    const uint64_t g = gbase[d] + i;
    out_h[g] = h;
    out_p[g] = tp[i];
    if (bounds && (i == dstart[d] || mulhi64(th[i - 1], count) != seg))
      atomicMin(bounds + seg, (uint32_t)g);
  }
}

// bounds[s] = *n_ptr for s <= count, before the last pass lowers them.
__global__ void bounds_init_kernel(uint32_t *bounds, uint64_t count, const uint64_t *n_ptr) {
  const uint32_t n = (uint32_t)*n_ptr;
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s <= count;
       s += (uint64_t)gridDim.x * blockDim.x)
    bounds[s] = n;
}

// One block: bounds[s] = min(bounds[s], bounds[s + 1], ...), so an empty
// segment starts where the next non-empty one does. Each thread takes a
// contiguous piece, then the pieces' minima are combined from the end.
__global__ __launch_bounds__(1024) void bounds_fill_kernel(uint32_t *bounds, uint64_t count) {
  __shared__ uint32_t piece_min[1024];
  const uint64_t n = count + 1, per = (n + 1023) / 1024;
  const uint64_t a = threadIdx.x * per, e = a + per < n ? a + per : n;
  uint32_t m = ~0u;
  for (uint64_t s = e; s-- > a;) {
    m = bounds[s] < m ? bounds[s] : m;
    bounds[s] = m;
  }
  piece_min[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int t = 1022; t >= 0; t--)
      piece_min[t] = piece_min[t + 1] < piece_min[t] ? piece_min[t + 1] : piece_min[t];
  __syncthreads();
  const uint32_t later = threadIdx.x < 1023 ? piece_min[threadIdx.x + 1] : ~0u;
  for (uint64_t s = a; s < e; s++)
    if (later < bounds[s]) bounds[s] = later;
}

// start[s] = first entry of segment s (start[count] = n).
__global__ void bucket_bounds_kernel(const uint64_t *h, const uint64_t *n_ptr, uint64_t count,
                                     uint64_t max_n, uint32_t *start) {
  const uint64_t n = *n_ptr;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= max_n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i > n) break;
    const int64_t sp = i ? (int64_t)mulhi64(h[i - 1], count) : -1;
    const int64_t sc = i < n ? (int64_t)mulhi64(h[i], count) : (int64_t)count;
    for (int64_t s = sp + 1; s <= sc; s++) start[s] = (uint32_t)i;
  }
}

// Member words of every hit row to zero (merged groups OR their bits in).
__global__ void zero_members_kernel(uint32_t *hits, uint64_t n_pos, uint32_t row_words,
                                    uint32_t first, uint32_t words) {
  const uint64_t total = n_pos * words;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x)
    hits[(i / words) * row_words + first + i % words] = 0;
}

// Hit rows of the chunk's positions that hold no entry (read separators,
// windows with invalid bases) to zero, so the sums read every row: one wave
// per 64 positions of the bitmap of listed positions.
__global__ __launch_bounds__(256) void zero_unlisted_kernel(uint32_t *hits, uint32_t row_words,
                                                            const uint64_t *listed,
                                                            uint64_t n_pos) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t words = (n_pos + 63) / 64;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t wi = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); wi < words;
       wi += waves) {
    uint64_t m = ~listed[wi];
    if (wi == words - 1 && (n_pos & 63)) m &= (1ULL << (n_pos & 63)) - 1;
    while (m) {
      const uint64_t p = wi * 64 + (uint64_t)__builtin_ctzll(m);
      m &= m - 1;
      for (uint32_t w = lane; w < row_words; w += 64) hits[p * row_words + w] = 0;
    }
  }
}

// Hit bits of the lane's 8 plane bytes (w0 low, w1 high), fields of b bits
// little endian (the plane layout, storage.rs:1721-1754): field f equal to
// fp sets bit field_bit<B>(f). At b = 8 the zero bytes of the two words are
// gathered by one multiply, which leaves the fields in the order
// f4 f0 f5 f1 f6 f2 f7 f3; every other width keeps field order.
template <int B>
__host__ __device__ __forceinline__ uint32_t field_bit(uint32_t f) {
  if constexpr (B == 8) return f < 4 ? 2 * f + 1 : 2 * (f - 4);
  else return f;
}
template <int B>
__device__ __forceinline__ uint32_t bit_field(uint32_t bit) {
  if constexpr (B == 8) return (bit & 1) ? (bit >> 1) : 4 + (bit >> 1);
  else return bit;
}

// b = 8: bit field_bit<8>(f) set where byte f of (v0, v1) is zero.
__device__ __forceinline__ uint64_t zero_bytes_to_bits(uint32_t v0, uint32_t v1) {
  // 0x80 in each zero byte
  const uint32_t z0 = ~(((v0 & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v0) & 0x80808080u;
  const uint32_t z1 = ~(((v1 & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v1) & 0x80808080u;
  // byte i's bits at 8i+6 (word 1, field 4+i) and 8i+7 (word 0, field i);
  // y * (2^18 + 2^12 + 2^6 + 1) moves them to bits 24+2i and 25+2i, and
  // no two partial products overlap, so nothing carries into bits 24..31
  const uint32_t y = z0 | (z1 >> 1);
  return __umulhi(y, 0x04104100u) & 0xFFu;  // (y * M) >> 24, bits 24..31
}

template <int B>
__device__ __forceinline__ uint64_t slice_hits(uint32_t w0, uint32_t w1, uint32_t fp) {
  if constexpr (B == 8) {
    const uint32_t fpb = fp * 0x01010101u;
    return zero_bytes_to_bits(w0 ^ fpb, w1 ^ fpb);
  } else if constexpr (B == 16) {
    auto two = [&](uint32_t w) {
      return (uint32_t)((w & 0xFFFFu) == fp) | ((uint32_t)((w >> 16) == fp) << 1);
    };
    return two(w0) | (two(w1) << 2);
  } else if constexpr (B == 32) {
    return (uint64_t)(w0 == fp) | ((uint64_t)(w1 == fp) << 1);
  } else if constexpr (B == 4) {
    auto eight = [&](uint32_t w) {
      const uint32_t v = w ^ (fp * 0x11111111u);
      uint32_t z = (~(((v & 0x77777777u) + 0x77777777u) | v) & 0x88888888u) >> 3;  // 0,4,..,28
      z = (z | (z >> 3)) & 0x03030303u;
      z = (z | (z >> 6)) & 0x000F000Fu;
      return (z | (z >> 12)) & 0xFFu;
    };
    return eight(w0) | (eight(w1) << 8);
  } else if constexpr (B == 2) {
    auto sixteen = [&](uint32_t w) {
      const uint32_t v = w ^ (fp * 0x55555555u);
      uint32_t z = ~(v | (v >> 1)) & 0x55555555u;  // even bits
      z = (z | (z >> 1)) & 0x33333333u;
      z = (z | (z >> 2)) & 0x0F0F0F0Fu;
      z = (z | (z >> 4)) & 0x00FF00FFu;
      return (z | (z >> 8)) & 0xFFFFu;
    };
    return (uint64_t)sixteen(w0) | ((uint64_t)sixteen(w1) << 16);
  } else {
    const uint64_t x = (uint64_t)w0 | ((uint64_t)w1 << 32);
    return fp ? x : ~x;
  }
}

// Bits of the fields f < n (n clamped to [0, F]) in field_bit order.
template <int B>
__device__ __forceinline__ uint64_t fields_below(int64_t n) {
  constexpr int F = 64 / B;
  if (n <= 0) return 0;
  if (n >= F) return F >= 64 ? ~0ULL : ((1ULL << F) - 1);
  if constexpr (B == 8) {
    uint64_t m = 0;
    for (int f = 0; f < n; f++) m |= 1ULL << field_bit<8>((uint32_t)f);
    return m;
  } else {
    return (1ULL << n) - 1;
  }
}

// Hit bits of one entry per lane over a 512-byte slice at b = 8: lane l's
// entry reads its H window rows (byte offsets rb[i], multiples of 512) in
// 16-byte pieces, piece (l ^ c) & 31 at step c, so the 16 lanes of each LDS
// access group read 16 different bank groups. Each piece holds two 8-byte
// chunks (8 fields each); the entry's 64 hit bytes are built in that order
// in o[] (half-word c = piece (l ^ c) & 31) and put back in field order.
template <int HT>
__device__ __forceinline__ void lane_slice_hits_b8(const uint8_t *win, uint32_t lane,
                                                   const uint32_t (&rb)[8], uint32_t H,
                                                   uint32_t fp, uint32_t (&o)[16]) {
  constexpr int HMAX = HT ? HT : 8;
  const uint32_t fpb = fp * 0x01010101u;
  const uint32_t t = lane & 31;
  uint32_t rl[HMAX];
#pragma unroll
  for (int i = 0; i < HMAX; i++) rl[i] = rb[i] ^ (t * 16);
#pragma unroll
  for (int c = 0; c < 32; c++) {
    uint4 r[HMAX];
#pragma unroll
    for (int i = 0; i < HMAX; i++)
      if (HT || i < (int)H) r[i] = *(const uint4 *)(win + (rl[i] ^ (uint32_t)(c * 16)));
    uint32_t a0, a1, b0, b1;
    if constexpr (HT == 4) {
      a0 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(r[0].x, r[1].x, r[2].x, 0x96), r[3].x, fpb, 0x96);
      a1 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(r[0].y, r[1].y, r[2].y, 0x96), r[3].y, fpb, 0x96);
      b0 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(r[0].z, r[1].z, r[2].z, 0x96), r[3].z, fpb, 0x96);
      b1 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(r[0].w, r[1].w, r[2].w, 0x96), r[3].w, fpb, 0x96);
    } else {
      a0 = a1 = b0 = b1 = fpb;
#pragma unroll
      for (int i = 0; i < HMAX; i++)
        if (HT || i < (int)H) {
          a0 ^= r[i].x;
          a1 ^= r[i].y;
          b0 ^= r[i].z;
          b1 ^= r[i].w;
        }
    }
    const uint32_t hw = (zero_bytes_to_bits(a0, a1) & 0xFFu) | ((zero_bytes_to_bits(b0, b1) & 0xFFu) << 8);
    if (c & 1) o[c >> 1] |= hw << 16;
    else o[c >> 1] = hw;
  }
  // field order: half-word k of the result is half-word k ^ t of o
  if (t & 1) {
#pragma unroll
    for (int d = 0; d < 16; d++) o[d] = (o[d] >> 16) | (o[d] << 16);
  }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const bool sw = (t >> (b + 1)) & 1;
#pragma unroll
    for (int d = 0; d < 16; d++)
      if (!(d & (1 << b))) {
        const uint32_t x = o[d], y = o[d | (1 << b)];
        o[d] = sw ? y : x;
        o[d | (1 << b)] = sw ? x : y;
      }
  }
}

}  // namespace

struct StageArgs {
  BankDev bank;
  RecDev rec;
  const uint64_t *ent_h;
  const uint32_t *ent_p;
  const uint32_t *start;  // [count + 1]
  uint64_t seed;
  uint32_t slice_bytes;   // LDS row stride: plane bytes per slice (<= 512)
  uint32_t n_slices;
  uint32_t *hits;         // entries' rows: positions (or keys) x row_words
  uint32_t row_words;
  uint32_t col_word0;     // first word of this bank's column bits in a row
  uint32_t member_bit0;   // first bit of this bank's member bits in a row
  uint32_t *queue;        // per-XCD unit counters at stride 32
  uint32_t xcds;
  unsigned long long *folded;  // entries answered from the entry before them
  const uint8_t *planes;  // the bank as bit planes (plane_lookup_kernel), or null
  uint64_t planes_pitch;  // its row pitch
  uint32_t parts;         // plane_lookup_kernel: entry ranges per (segment, slice)
  uint32_t row_stage;     // plane_lookup_kernel: hit rows through LDS, whole per store
};

namespace {

// HT: the hash count when known at compile time (2, 3, 4), 0 for any H <= 8.
template <int B, int HT>
__global__ __launch_bounds__(1024) void stage_lookup_kernel(StageArgs a) {
  constexpr int F = 64 / B;          // fields per lane (8 bytes)
  constexpr uint32_t OUT = 8 * F;    // hit bytes per slice (64 lanes x F bits)
  constexpr int HMAX = HT ? HT : 8;
  __shared__ __attribute__((aligned(16))) uint8_t win[kStageWindow];
  __shared__ unsigned long long rec_bits_all[kStageWaves * 64];
  // each wave's hit bytes of G entries, written out together
  constexpr uint32_t G = OUT >= kStageOut ? 1 : (kStageOut / OUT < 8 ? kStageOut / OUT : 8);
  __shared__ __attribute__((aligned(16))) uint8_t out_all[kStageWaves * kStageOut];
  __shared__ uint32_t unit_sh;
  const BankDev &bk = a.bank;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint8_t *staged = out_all + wave * kStageOut;
  unsigned long long *rec_bits = rec_bits_all + wave * 64;
  const uint32_t H = HT ? (uint32_t)HT : bk.hashes;
  const uint64_t L = bk.geo.length, count = bk.geo.count;
  const uint32_t SB = a.slice_bytes;
  const uint32_t lane_off = lane * 8 < SB ? lane * 8 : SB - 8;  // lanes past the plane are masked
  const uint32_t x = blockIdx.x % a.xcds;
  const uint64_t s_lo = count * x / a.xcds, s_hi = count * (x + 1) / a.xcds;
  const uint64_t units = (s_hi - s_lo) * a.n_slices;
  const uint64_t seed = a.seed;
  uint64_t folded = 0;
  while (true) {
    __syncthreads();  // the previous unit's window is no longer read
    if (threadIdx.x == 0) unit_sh = atomicAdd(a.queue + 32 * x, 1u);
    __syncthreads();
    const uint64_t u = unit_sh;
    if (u >= units) break;
    const uint64_t s = s_lo + u / a.n_slices;
    const uint32_t q = (uint32_t)(u % a.n_slices);
    const uint32_t e0 = a.start[s], e1 = a.start[s + 1];
    if (e0 == e1) continue;
    // the window: rows s*L .. (s+H)*L - 1, bytes [q*SB, (q+1)*SB) of each,
    // 16-byte pieces straight from memory into LDS: a wave's instruction
    // fills 64 consecutive pieces, piece c at LDS offset 16c (the window's
    // row stride SB is per_row pieces)
    {
      const uint8_t *src = bk.data + s * L * bk.row_bytes + (uint64_t)q * SB;
      const uint32_t per_row = SB / 16, chunks = H * (uint32_t)L * per_row;
      for (uint32_t cb = wave * 64; cb < chunks; cb += kStageThreads) {
        const uint32_t c = cb + lane;
        if (c < chunks)
          __builtin_amdgcn_global_load_lds(
              (const void *)(src + (uint64_t)(c / per_row) * bk.row_bytes + (c % per_row) * 16),
              (__attribute__((address_space(3))) void *)(win + cb * 16), 16, 0, 0);
      }
    }
    // this wave's first entries load while the window fills; each batch then
    // loads the next one's before its own lookups
    uint64_t h_next = 0;
    uint32_t p_next = 0;
    if (e0 + wave * 64 + lane < e1) {
      h_next = a.ent_h[e0 + wave * 64 + lane];
      p_next = a.ent_p[e0 + wave * 64 + lane];
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    // this lane's fields: columns c0 .. c0 + F - 1
    const int64_t c0 = (int64_t)q * SB * 8 / B + (int64_t)lane * F;
    const uint64_t fmask = fields_below<B>((int64_t)bk.groups - c0);
    const uint64_t mmask = fmask & ~fields_below<B>((int64_t)bk.first_merged - c0);
    const bool all_fields = __ballot(fmask != fields_below<B>(F)) == 0;
    const bool any_merged = __ballot(mmask != 0) != 0;  // merged groups in this slice
    const uint32_t slice_c0 = q * SB * 8 / B, slice_c1 = slice_c0 + 64 * F;
    // b = 8, one entry per lane: the valid fields of a slice's 16 hit words
    // (the slice holding the last groups has fewer than 512)
    uint32_t tail_mask[16];
    if constexpr (B == 8 && HT != 0) {
      const int32_t nv = (int32_t)std::min<int64_t>(std::max<int64_t>((int64_t)bk.groups - slice_c0, 0), 512);
      const int32_t full = nv >> 3;  // whole valid 8-field chunks
      const uint32_t part = (uint32_t)fields_below<8>(nv & 7);
#pragma unroll
      for (int d = 0; d < 16; d++) {
        uint32_t mk = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int32_t c = 4 * d + k;
          mk |= (c < full ? 0xFFu : c == full ? part : 0u) << (8 * k);
        }
        tail_mask[d] = mk;
      }
    }
    for (uint32_t b0 = e0 + wave * 64; b0 < e1; b0 += kStageWaves * 64) {
      const uint32_t e = b0 + lane;
      const bool v = e < e1;
      const uint64_t h = h_next;
      const uint32_t p = p_next;
      if (e + kStageWaves * 64 < e1) {
        h_next = a.ent_h[e + kStageWaves * 64];
        p_next = a.ent_p[e + kStageWaves * 64];
      }
      const uint64_t key = unsplitmix64(h) - seed;  // mixsplit(key, seed) == h
      const uint32_t fp = (uint32_t)fingerprint(key, seed, B);
      const uint64_t step = splitmix64(h);
      // window rows of the H hashes, i*L + ((h_i ^ h_i >> 33) & (L - 1)), as
      // byte offsets in the window in 64-byte units (SB is a multiple of 64,
      // the window at most 128 KiB), 16 bits each: hashes 0..3 in offs,
      // 4..7 in offs_hi
      const uint32_t sb64 = SB / 64;
      uint64_t offs = 0, offs_hi = 0, hv = h;
#pragma unroll
      for (int i = 0; i < (HMAX < 4 ? HMAX : 4); i++)
        if (HT || i < (int)H) {
          offs |= (uint64_t)((i * L + ((hv ^ (hv >> 33)) & (L - 1))) * sb64) << (16 * i);
          hv += step;
        }
      if (!HT && H > 4)
        for (uint32_t i = 4; i < H && i < 8; i++) {
          offs_hi |= (uint64_t)((i * L + ((hv ^ (hv >> 33)) & (L - 1))) * sb64) << (16 * (i - 4));
          hv += step;
        }
      // row offsets and the fingerprint are computed here, before the entry
      // loop, so no load is outstanding when the loop reads them
      const uint64_t prow = (uint64_t)p * a.row_words;
      // recovery directory probe (open addressing on the high hash bits); its
      // first load is in flight during the lane lookup below
      const bool probe = a.rec.enabled && v;
      const uint32_t ah = probe ? abandoned_hash(key, seed) : 0;
      const uint64_t cap_mask = (1ULL << a.rec.cap_log2) - 1;
      uint64_t slot = a.rec.cap_log2 ? (uint64_t)(ah >> (32 - a.rec.cap_log2)) : 0;
      uint64_t de = probe ? a.rec.dir[slot] : (uint64_t)kNone;
      // a partitioned bank answers the keys routed to it (storage.rs:99-101)
      const uint64_t own_mask =
          __ballot(v && (bk.partitions <= 1 || partition_id(key, seed, bk.partitions) == bk.partition));
      // b = 8 over whole 512-byte slices with no merged group: one entry per
      // lane, its row written once the probe has ended
      bool by_lane = false;
      uint32_t o[16];
      if constexpr (B == 8 && HT != 0) {
        if (SB == 512 && !any_merged) {
          by_lane = true;
          if ((own_mask >> lane) & 1) {
            uint32_t rb[8];
#pragma unroll
            for (int i = 0; i < 4; i++) {
              rb[i] = (uint32_t)((offs >> (16 * i)) & 0xFFFFu) * 64;
              rb[i + 4] = (uint32_t)((offs_hi >> (16 * i)) & 0xFFFFu) * 64;
            }
            lane_slice_hits_b8<HT>(win, lane, rb, H, fp, o);
            if (!all_fields) {
#pragma unroll
              for (int d = 0; d < 16; d++) o[d] &= tail_mask[d];
            }
          }
        }
      }
      uint32_t rec_start = 0, rec_len = 0;
      if (probe) {
        while (true) {
          const uint32_t pi = (uint32_t)de;
          if (pi == kNone) break;
          if ((uint32_t)(de >> 32) == ah) {
            rec_len = a.rec.post[pi];
            rec_start = pi + 1;
            break;
          }
          slot = (slot + 1) & cap_mask;
          de = a.rec.dir[slot];
        }
      }
      const uint64_t rec_mask = __ballot(rec_len != 0);
      if constexpr (B == 8 && HT != 0) {
        if (by_lane) {
          if ((own_mask >> lane) & 1) {
            uint4 *dst = (uint4 *)((uint8_t *)(a.hits + prow + a.col_word0) + (uint64_t)q * 64);
#pragma unroll
            for (int k = 0; k < 4; k++)
              dst[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
          }
          const uint64_t fix = own_mask & rec_mask;
          if (fix) {
            // recovery (storage.rs:2771-2787, 2916-2924): a posting's colour
            // counts as a hit. The rows are in memory before the bits of
            // the recovered colours in this slice are ORed in; the postings
            // of all recovered entries of the batch go 64 per pass.
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            const uint32_t len = ((fix >> lane) & 1) ? rec_len : 0u;
            const uint32_t end = (uint32_t)wave_inclusive_scan(len, lane);
            const uint32_t total = (uint32_t)__shfl((int)end, 63, 64);
            uint32_t *ends = (uint32_t *)rec_bits;  // inclusive ends of the lanes' postings
            ends[lane] = end;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t g0 = 0; g0 < total; g0 += 64) {
              const uint32_t g = g0 + lane < total ? g0 + lane : total - 1;
              uint32_t j = 0;  // the first lane whose postings end after g
#pragma unroll
              for (uint32_t st = 32; st; st >>= 1)
                if (ends[j + st - 1] <= g) j += st;
              const uint32_t first = j ? ends[j - 1] : 0u;
              const uint32_t rs = (uint32_t)__shfl((int)rec_start, (int)j, 64);
              const uint64_t pj = shfl64(prow, j);
              if (g0 + lane < total) {
                const uint32_t pe = a.rec.post_col[rs + (g - first)];
                const uint32_t col = pe & 0xFFFFFFu;
                if (pe != kNone && (pe >> 24) == bk.colmap && col >= slice_c0 && col < slice_c1 &&
                    col < bk.groups) {
                  const uint32_t rel = col - slice_c0, ch = rel / 8;
                  atomicOr(a.hits + pj + a.col_word0 + q * 16 + ch / 4,
                           1u << (8 * (ch % 4) + field_bit<8>(rel % 8)));
                }
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          }
        }
      }
      // an entry whose seed equals the entry before it reuses its hits
      const uint64_t prev_h = shfl_up64(h, 1);
      const uint64_t same_mask =
          by_lane ? 0 : __ballot(v && lane > 0 && prev_h == h) & own_mask & (own_mask << 1);
      folded += __popcll(same_mask);
      uint64_t m_seed = 0;
      // staged entries (their lanes j) and the write-out: lane l stores 16
      // bytes of entry l / (OUT / 16), so a row's OUT bytes are one access
      uint64_t staged_js = 0;  // lane j of the entry in slot t: bits 8t .. 8t+5
      uint32_t n_staged = 0;
      auto flush_staged = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        constexpr uint32_t per = OUT / 16;  // lanes per entry (OUT >= 16)
        const uint32_t r = lane / per;
        const uint32_t jr = r < G ? (uint32_t)(staged_js >> (8 * r)) & 63u : 0u;
        const uint64_t pr = shfl64(prow, jr);
        if (r < n_staged) {
          uint8_t *dst = (uint8_t *)(a.hits + pr + a.col_word0) + (uint64_t)q * OUT;
          const uint32_t c = (lane % per) * 16;
          *(uint4 *)(dst + c) = *(const uint4 *)(staged + r * OUT + c);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        staged_js = 0;
      };
      const uint8_t *lane_win = win + lane_off;
      // The H rows of entry j XORed and compared in this slice (from LDS).
      auto seed_hits = [&](uint32_t j) -> uint64_t {
        const uint64_t oj = readlane64(offs, (int)j);
        const uint32_t fpj = readlane32(fp, (int)j);
        uint32_t w0 = 0, w1 = 0;
        uint2 rv[HMAX < 4 ? HMAX : 4];
#pragma unroll
        for (int i = 0; i < (HMAX < 4 ? HMAX : 4); i++) {
          const uint32_t o64 = (uint32_t)(oj >> (16 * i)) & 0xFFFFu;
          rv[i] = (HT || i < (int)H) ? *(const uint2 *)(lane_win + o64 * 64) : make_uint2(0, 0);
        }
        if constexpr (B == 8 && (HT == 3 || HT == 4)) {
          // rows and the fingerprint in 3-input XORs; fp is folded in here
          const uint32_t fpb = fpj * 0x01010101u;
          if constexpr (HT == 4) {
            w0 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(rv[0].x, rv[1].x, rv[2].x, 0x96),
                                             rv[3].x, fpb, 0x96);
            w1 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(rv[0].y, rv[1].y, rv[2].y, 0x96),
                                             rv[3].y, fpb, 0x96);
          } else {
            w0 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(rv[0].x, rv[1].x, rv[2].x, 0x96),
                                             fpb, 0u, 0x96);
            w1 = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(rv[0].y, rv[1].y, rv[2].y, 0x96),
                                             fpb, 0u, 0x96);
          }
          uint64_t mm = zero_bytes_to_bits(w0, w1);
          if (!all_fields) mm &= fmask;
          return mm;
        }
#pragma unroll
        for (int i = 0; i < (HMAX < 4 ? HMAX : 4); i++) {
          w0 ^= rv[i].x;
          w1 ^= rv[i].y;
        }
        if (!HT && H > 4) {
          const uint64_t ohj = readlane64(offs_hi, (int)j);
          for (uint32_t i = 4; i < H; i++) {
            const uint32_t o64 = (uint32_t)(ohj >> (16 * (i - 4))) & 0xFFFFu;
            const uint2 r = *(const uint2 *)(lane_win + o64 * 64);
            w0 ^= r.x;
            w1 ^= r.y;
          }
        }
        uint64_t mm = slice_hits<B>(w0, w1, fpj);
        if (!all_fields) mm &= fmask;
        return mm;
      };
      // Hot loop: entries without recovery postings, lowest lane first (a
      // scalar find-first-set); it reads only LDS and writes only staged
      // rows. Entries whose merged groups hit are fixed up below.
      uint64_t fix_mask = 0;
      for (uint64_t todo = by_lane ? 0 : own_mask & ~rec_mask; todo; todo &= todo - 1) {
        const uint32_t j = (uint32_t)__builtin_ctzll(todo);
        if (!((same_mask >> j) & 1)) m_seed = seed_hits(j);
        uint64_t m = m_seed;
        if (any_merged) {
          if (__ballot((m & mmask) != 0)) fix_mask |= 1ULL << j;
          m &= ~mmask;  // a merged group's own bit is no colour
        }
        // this entry's OUT hit bytes go to slot n_staged of the wave's buffer
        uint8_t *out = staged + n_staged * OUT;
        if constexpr (F == 64) ((uint64_t *)out)[lane] = m;
        else if constexpr (F == 32) ((uint32_t *)out)[lane] = (uint32_t)m;
        else if constexpr (F == 16) ((uint16_t *)out)[lane] = (uint16_t)m;
        else if constexpr (F == 8) out[lane] = (uint8_t)m;
        else if constexpr (F == 4) {
          const uint32_t up = __shfl_down((uint32_t)m, 1);
          if (!(lane & 1)) out[lane >> 1] = (uint8_t)((m & 0xFu) | ((up & 0xFu) << 4));
        } else {  // F == 2: four lanes per byte
          uint32_t bits2 = (uint32_t)m & 3u;
          bits2 |= (__shfl_down(bits2, 1) & 3u) << 2;
          bits2 |= (__shfl_down(bits2, 2) & 15u) << 4;
          if (!(lane & 3)) out[lane >> 2] = (uint8_t)bits2;
        }
        staged_js |= (uint64_t)j << (8 * n_staged);
        if (++n_staged == G) {
          flush_staged();
          n_staged = 0;
        }
      }
      if (n_staged) {
        flush_staged();
        n_staged = 0;
      }
      // Rare pass: entries with recovery postings (written here whole) and
      // entries whose merged groups hit (member bits ORed into their rows).
      for (uint64_t todo = by_lane ? 0 : (own_mask & rec_mask) | fix_mask; todo; todo &= todo - 1) {
        const uint32_t j = (uint32_t)__builtin_ctzll(todo);
        const uint64_t ms = seed_hits(j);
        uint64_t m = ms;
        uint32_t *row_out = a.hits + readlane64(prow, (int)j);
        const bool rec = (rec_mask >> j) & 1;
        // recovery: a recovered colour counts as a hit (storage.rs:2771-2787,
        // 2916-2924); postings spread over the lanes, 64 per pass
        if (rec) {
          const uint32_t rl = readlane32(rec_len, (int)j);
          const uint32_t rs = readlane32(rec_start, (int)j);
          rec_bits[lane] = 0;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          for (uint32_t q0 = 0; q0 < rl; q0 += 64) {
            const uint32_t qq = q0 + lane;
            const uint32_t pe = qq < rl ? a.rec.post_col[rs + qq] : kNone;
            const uint32_t col = pe & 0xFFFFFFu;
            const bool mine = pe != kNone && (pe >> 24) == bk.colmap && col >= slice_c0 &&
                              col < slice_c1 && col < bk.groups;
            if (mine && col < bk.first_merged)
              atomicOr(&rec_bits[(col - slice_c0) / F],
                       1ULL << field_bit<B>((col - slice_c0) % F));
            if (mine && col >= bk.first_merged) {
              const uint32_t o = a.rec.post[rs + qq];
              const uint32_t bit =
                  a.member_bit0 + bk.mg_off[col - bk.first_merged] + bk.color_bit[o];
              atomicOr(&row_out[bit >> 5], 1u << (bit & 31));
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          m |= rec_bits[lane] & fmask;
        }
        if (any_merged) {
          // merged groups: decode the XORed membership mask (get_bits on the
          // H rows) and set each member's bit
          const uint64_t kj = readlane64(key, (int)j);
          uint64_t mh = ms & mmask;
          while (mh) {
            const uint32_t f = bit_field<B>((uint32_t)__builtin_ctzll(mh));
            mh &= mh - 1;
            const uint32_t g = (uint32_t)(c0 + f) - bk.first_merged;
            const uint32_t lo = bk.mg_off[g], width = bk.mg_off[g + 1] - lo;
            uint64_t mask = xor_rows_bits(bk, kj, seed, bk.mg_mask_bit[g], width);
            while (mask) {
              const uint32_t bit = a.member_bit0 + lo + __builtin_ctzll(mask);
              mask &= mask - 1;
              atomicOr(&row_out[bit >> 5], 1u << (bit & 31));
            }
          }
          m &= ~mmask;
        }
        if (rec) {  // the hot loop skipped this entry: its column bytes
          uint8_t *out = (uint8_t *)(row_out + a.col_word0) + (uint64_t)q * OUT;
          if constexpr (F == 64) ((uint64_t *)out)[lane] = m;
          else if constexpr (F == 32) ((uint32_t *)out)[lane] = (uint32_t)m;
          else if constexpr (F == 16) ((uint16_t *)out)[lane] = (uint16_t)m;
          else if constexpr (F == 8) out[lane] = (uint8_t)m;
          else if constexpr (F == 4) {
            const uint32_t up = __shfl_down((uint32_t)m, 1);
            if (!(lane & 1)) out[lane >> 1] = (uint8_t)((m & 0xFu) | ((up & 0xFu) << 4));
          } else {
            uint32_t bits2 = (uint32_t)m & 3u;
            bits2 |= (__shfl_down(bits2, 1) & 3u) << 2;
            bits2 |= (__shfl_down(bits2, 2) & 15u) << 4;
            if (!(lane & 3)) out[lane >> 2] = (uint8_t)bits2;
          }
        }
      }
    }
  }
  if (folded && lane == 0) atomicAdd(a.folded, (unsigned long long)folded);
}

// Field c (B bits from bit c * B, little endian) of a plane-layout row.
template <int B>
__device__ __forceinline__ uint32_t plane_field(const uint8_t *row, uint32_t c) {
  if constexpr (B >= 8) {
    const uint8_t *p = row + (uint64_t)c * (B / 8);
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < B / 8; i++) v |= (uint32_t)p[i] << (8 * i);
    return v;
  } else {
    const uint64_t bit = (uint64_t)c * B;
    return (row[bit >> 3] >> (bit & 7)) & ((1u << B) - 1);
  }
}

// The bank as bit planes: row r of `out` (pitch bytes) holds the row's
// fingerprint fields in groups of 32; group g, at byte 4 * B * g, is B words,
// bit f of word k being bit k of field 32 g + f. Fields past the last column
// stay zero (the rows are cleared before).
template <int B>
__global__ void bit_planes_kernel(const uint8_t *data, uint64_t row_bytes, uint64_t rows,
                                  uint32_t groups, uint32_t g32, uint64_t pitch, uint8_t *out) {
  const uint64_t total = rows * g32;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = i / g32;
    const uint32_t g = (uint32_t)(i % g32);
    const uint8_t *row = data + r * row_bytes;
    uint32_t pl[B];
#pragma unroll
    for (int k = 0; k < B; k++) pl[k] = 0;
    for (uint32_t f = 0; f < 32; f++) {
      const uint32_t c = g * 32 + f;
      if (c >= groups) break;
      const uint32_t v = plane_field<B>(row, c);
#pragma unroll
      for (int k = 0; k < B; k++) pl[k] |= ((v >> k) & 1u) << f;
    }
    uint32_t *dst = (uint32_t *)(out + r * pitch) + (uint64_t)g * B;
#pragma unroll
    for (int k = 0; k < B; k++) dst[k] = pl[k];
  }
}

// XOR of the H row words of one plane word and the fingerprint bit mask fm.
template <int HT, int HMAX>
__device__ __forceinline__ uint32_t rows_xor(const uint32_t (&w)[HMAX], uint32_t fm, uint32_t H) {
  if constexpr (HT == 4) {
    return __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(w[0], w[1], w[2], 0x96), w[3], fm,
                                       0x96);
  } else if constexpr (HT == 3) {
    return __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(w[0], w[1], w[2], 0x96), fm, 0u,
                                       0x96);
  } else if constexpr (HT == 2) {
    return __builtin_amdgcn_bitop3_b32(w[0], w[1], fm, 0x96);
  } else {
    uint32_t d = fm;
#pragma unroll
    for (int i = 0; i < HMAX; i++)
      if (i < (int)H) d ^= w[i];
    return d;
  }
}

// LDS written by a wave visible to its own later LDS reads (and the reverse)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bytes of LDS per wave through which plane_lookup_kernel passes its hit rows
constexpr uint32_t kRowStageBytes = 2048;

// The staged lookup over the bank as bit planes (bit_planes_kernel), for
// banks without merged groups: per 32 fields, the H rows' B plane words are
// XORed with the fingerprint's bits (all-ones or all-zero words) and ORed, so
// a field hits where no plane differs -- about 5 instructions per 8 fields
// at b = 8 and H = 4. A block (8 waves, two blocks per CU) takes units
// (segment, slice of NP 16-byte pieces per row) from its XCD's counter and
// copies the slice of the segment's H x L window rows into LDS. Each lane
// takes one entry and reads its rows' pieces in the order (l ^ c) mod NP, so
// the 16 lanes of an LDS access read 16 different bank groups. Hit words
// are written in field order: bit f of word w of slice q is column
// 32 (q SG + w) + f. A lane's hit words for the slice (SG words) go to
// LDS, R entries at a time, and are read back 16 bytes per lane so that each
// store instruction writes 64 / NQ whole row slices: random rows written
// whole by one instruction are stored several times faster than 16-byte
// pieces of 64 rows. Recovery postings are ORed in afterwards, once per
// entry (plane_recovery_kernel).
template <int B, int HT, int NP>
__global__ __launch_bounds__(plane_threads(B)) void plane_lookup_kernel(StageArgs a) {
  constexpr uint32_t kPlaneWaves = plane_threads(B) / 64;
  constexpr uint32_t SB = NP * 16;              // bytes of a row slice
  constexpr uint32_t SG = SB / (4 * B);         // 32-field groups (hit words) per slice
  constexpr uint32_t PPG = B >= 4 ? B / 4 : 1;  // pieces per group
  constexpr uint32_t GPP = B >= 4 ? 1 : 4 / B;  // groups per piece
  constexpr uint32_t BS = GPP;                  // hit words per block of the piece order
  constexpr uint32_t NB = SG / BS;
  constexpr int HMAX = HT ? HT : 8;
  static_assert(SG >= 1 && NB >= 1 && (NP & (NP - 1)) == 0, "slice geometry");
  // hit rows through LDS: NQ 16-byte pieces per entry, R entries per round
  constexpr uint32_t NQ = SG % 4 == 0 ? SG / 4 : 1;
  constexpr bool kRows = NQ >= 2;
  constexpr uint32_t R = kRowStageBytes / (16 * NQ) < 64 ? kRowStageBytes / (16 * NQ) : 64;
  static_assert(!kRows || (R % NQ == 0 && 64 % R == 0 && (R * NQ) % 64 == 0), "row staging");
  __shared__ __attribute__((aligned(16))) uint8_t win[plane_window(B)];
  __shared__ __attribute__((aligned(16))) uint8_t rowst[kRows ? kPlaneWaves * kRowStageBytes : 16];
  // the unit number shares wave 0's staging: it is read between the two
  // barriers that open a unit, staging is written only after the third
  uint32_t &unit_sh = *(uint32_t *)rowst;
  const BankDev &bk = a.bank;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t H = HT ? (uint32_t)HT : bk.hashes;
  const uint64_t L = bk.geo.length, count = bk.geo.count;
  const uint32_t x = blockIdx.x % a.xcds;
  const uint64_t s_lo = count * x / a.xcds, s_hi = count * (x + 1) / a.xcds;
  const uint64_t units = (s_hi - s_lo) * a.parts * a.n_slices;
  const uint64_t seed = a.seed;
  const uint32_t lp = (lane & (NP - 1)) * 16;  // the lane's first piece
  // hit words come out in blocks of BS words, block j of the slice at
  // block j ^ tq of o[]
  const uint32_t tq = (B >= 4 ? lane / PPG : lane) & (NB - 1);
  // the planes each piece of a group holds for this lane: 4 ((l ^ t) mod PPG)
  // .. + 3 at run position t (b >= 4); planes 0, 1, 0, 1 (b = 2); plane 0 (b = 1)
  uint32_t psh[PPG];
#pragma unroll
  for (uint32_t t = 0; t < PPG; t++) psh[t] = B >= 4 ? 4 * ((lane ^ t) & (PPG - 1)) : 0;
  while (true) {
    __syncthreads();  // the previous unit's window is no longer read
    if (threadIdx.x == 0) unit_sh = atomicAdd(a.queue + 32 * x, 1u);
    __syncthreads();
    const uint64_t u = unit_sh;
    if (u >= units) break;
    // unit u: slice q of entry range `part` of segment s, the slices and
    // ranges of one segment on neighbouring units (so the pieces of each
    // hit row are written close together in time)
    const uint32_t q = (uint32_t)(u % a.n_slices);
    const uint64_t t = u / a.n_slices;
    const uint32_t part = (uint32_t)(t % a.parts);
    const uint64_t s = s_lo + t / a.parts;
    const uint32_t s0 = a.start[s], len = a.start[s + 1] - s0;
    const uint32_t e0 = s0 + (uint32_t)((uint64_t)len * part / a.parts);
    const uint32_t e1 = s0 + (uint32_t)((uint64_t)len * (part + 1) / a.parts);
    if (e0 == e1) continue;
    // the window: slice q of rows s*L .. (s+H)*L - 1, straight into LDS
    {
      const uint8_t *src = a.planes + s * L * a.planes_pitch + (uint64_t)q * SB;
      const uint32_t chunks = H * (uint32_t)L * NP;
      for (uint32_t cb = wave * 64; cb < chunks; cb += plane_threads(B)) {
        const uint32_t c = cb + lane;
        if (c < chunks)
          __builtin_amdgcn_global_load_lds(
              (const void *)(src + (uint64_t)(c / NP) * a.planes_pitch + (c % NP) * 16),
              (__attribute__((address_space(3))) void *)(win + cb * 16), 16, 0, 0);
      }
    }
    uint64_t h_next = 0;
    uint32_t p_next = 0;
    if (e0 + wave * 64 + lane < e1) {
      h_next = a.ent_h[e0 + wave * 64 + lane];
      p_next = a.ent_p[e0 + wave * 64 + lane];
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (uint32_t b0 = e0 + wave * 64; b0 < e1; b0 += kPlaneWaves * 64) {
      const uint32_t e = b0 + lane;
      const bool v = e < e1;
      const uint64_t h = h_next;
      const uint32_t p = p_next;
      if (e + kPlaneWaves * 64 < e1) {
        h_next = a.ent_h[e + kPlaneWaves * 64];
        p_next = a.ent_p[e + kPlaneWaves * 64];
      }
      const uint64_t key = unsplitmix64(h) - seed;  // mixsplit(key, seed) == h
      const uint32_t fp = (uint32_t)fingerprint(key, seed, B);
      const uint64_t step = splitmix64(h);
      // window byte offsets of the H rows (i*L + ((h_i ^ h_i >> 33) & (L - 1)))
      // plus this lane's first piece
      uint32_t rb[HMAX];
      uint64_t hv = h;
#pragma unroll
      for (int i = 0; i < HMAX; i++) {
        rb[i] = 0;
        if (HT || i < (int)H) {
          rb[i] = (uint32_t)(i * L + ((hv ^ (hv >> 33)) & (L - 1))) * SB + lp;
          hv += step;
        }
      }
      const uint64_t prow = (uint64_t)p * a.row_words;
      // a partitioned bank answers the keys routed to it (storage.rs:99-101)
      const bool own =
          v && (bk.partitions <= 1 || partition_id(key, seed, bk.partitions) == bk.partition);
      const bool staged = kRows && a.row_stage;  // uniform
      uint32_t o[SG];
      if (!own) {
#pragma unroll
        for (uint32_t w = 0; w < SG; w++) o[w] = 0;
      } else {
        uint32_t fm[PPG][4];
#pragma unroll
        for (uint32_t t = 0; t < PPG; t++) {
          const uint32_t fpt = fp >> psh[t];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            fm[t][k] = 0u - ((fpt >> (B >= 4 ? k : (B == 2 ? (k & 1) : 0))) & 1u);
            asm volatile("" : "+v"(fm[t][k]));  // kept in a register across the pieces
          }
        }
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t c = 0; c < NP; c++) {
          uint4 r[HMAX];
#pragma unroll
          for (int i = 0; i < HMAX; i++)
            if (HT || i < (int)H) r[i] = *(const uint4 *)(win + (rb[i] ^ (c * 16)));
          const uint32_t t = c & (PPG - 1);
          uint32_t wx[HMAX], wy[HMAX], wz[HMAX], ww[HMAX];
#pragma unroll
          for (int i = 0; i < HMAX; i++) {
            wx[i] = r[i].x;
            wy[i] = r[i].y;
            wz[i] = r[i].z;
            ww[i] = r[i].w;
          }
          const uint32_t d0 = rows_xor<HT, HMAX>(wx, fm[t][0], H);
          const uint32_t d1 = rows_xor<HT, HMAX>(wy, fm[t][1], H);
          const uint32_t d2 = rows_xor<HT, HMAX>(wz, fm[t][2], H);
          const uint32_t d3 = rows_xor<HT, HMAX>(ww, fm[t][3], H);
          if constexpr (B >= 4) {
            // a field differs from the fingerprint where any plane does
            const uint32_t any = (t == 0 ? 0u : acc) | d0 | d1 | d2 | d3;
            if (t == PPG - 1) o[c / PPG] = ~any;
            else acc = any;
          } else if constexpr (B == 2) {
            o[2 * c] = ~(d0 | d1);
            o[2 * c + 1] = ~(d2 | d3);
          } else {
            o[4 * c] = ~d0;
            o[4 * c + 1] = ~d1;
            o[4 * c + 2] = ~d2;
            o[4 * c + 3] = ~d3;
          }
        }
        // into field order: block j <- block j ^ tq, i.e. word w <- word
        // w ^ tq BS. Staged rows take the part of the XOR above the low two
        // bits in the LDS slot of each 16-byte group instead.
#pragma unroll
        for (uint32_t sbit = 1; sbit < NB; sbit <<= 1) {
          if (staged && sbit * BS >= 4) continue;
          const bool sw = (tq & sbit) != 0;
#pragma unroll
          for (uint32_t j = 0; j < NB; j++)
            if (!(j & sbit)) {
#pragma unroll
              for (uint32_t w = 0; w < BS; w++) {
                const uint32_t x0 = o[j * BS + w], x1 = o[(j | sbit) * BS + w];
                o[j * BS + w] = sw ? x1 : x0;
                o[(j | sbit) * BS + w] = sw ? x0 : x1;
              }
            }
        }
        uint32_t *dst = a.hits + prow + a.col_word0 + (uint64_t)q * SG;
        if (!staged) {
          if constexpr (SG % 4 == 0) {
#pragma unroll
            for (uint32_t k = 0; k < SG / 4; k++)
              ((uint4 *)dst)[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
          } else if constexpr (SG == 2) {
            *(uint2 *)dst = make_uint2(o[0], o[1]);
          } else {
            dst[0] = o[0];
          }
        }
      }
      if constexpr (kRows) {
        if (staged) {
          // round r: entries (lanes) rR .. rR + R - 1 put their 16-byte
          // group k, which is piece k ^ (tq BS / 4) of the slice, at slot
          // k ^ (tq BS / 4) ^ (lane mod NQ) of their NQ slots; lane l then
          // reads slot 64i + l and stores it as piece (slot mod NQ) ^ (entry
          // mod NQ) of that entry's slice
          uint8_t *st = rowst + wave * kRowStageBytes;
          const uint32_t pv = own ? p : ~0u;
          const uint32_t xk = (((tq * BS) >> 2) ^ lane) % NQ;
          const uint32_t wb = (lane % R) * NQ * 16 + xk * 16;
#pragma unroll
          for (uint32_t r = 0; r < 64 / R; r++) {
            if (lane / R == r) {
#pragma unroll
              for (uint32_t k = 0; k < NQ; k++)
                *(uint4 *)(st + (wb ^ (k * 16))) =
                    make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
            }
            wave_lds_sync();
#pragma unroll
            for (uint32_t i = 0; i < R * NQ / 64; i++) {
              const uint32_t slot = i * 64 + lane, ent = slot / NQ;
              const uint4 val = *(const uint4 *)(st + slot * 16);
              const uint32_t pe = (uint32_t)__shfl((int)pv, (int)(r * R + ent), 64);
              const uint32_t k = (slot % NQ) ^ (ent % NQ);
              if (pe != ~0u)
                ((uint4 *)(a.hits + (uint64_t)pe * a.row_words + a.col_word0 +
                           (uint64_t)q * SG))[k] = val;
            }
            wave_lds_sync();
          }
        }
      }
    }
  }
}

// Recovery for the classes looked up as bit planes, once per entry after all
// of the class's lookups (so it follows every row store): an entry whose key
// has postings in the directory (storage.rs:2771-2787, 2916-2924) gets the
// bits of its postings' columns of this class ORed into its hit row. Keys
// with any of their bits clear in the directory's hash filter (128 KiB,
// staged in LDS by each persistent block) have no postings. The others queue up per wave
// in LDS and are probed 64 at a time, one per lane, so a wave waits on the
// directory once per 64 candidates rather than once per 64 entries. The
// postings of the recovered entries go 64 per pass, lane g finding its entry
// by binary search over the inclusive ends. Partitions of a class share its
// columns, so every entry is handled here whichever bank it routes to.
constexpr uint32_t kRecThreads = 1024;
constexpr uint32_t kRecQueue = 128;  // candidates queued per wave, at most

// Probes the directory for the first cnt (<= 64) queued candidates, one per
// lane, and ORs the postings of those that have some into their hit rows.
__device__ void recover_queued(const uint32_t *q_e, const uint32_t *q_ah, uint32_t cnt,
                               uint32_t *ends, const uint32_t *ent_p, const RecDev &rec,
                               uint32_t colmap, uint32_t groups, uint32_t *hits, uint32_t row_words,
                               uint32_t col_word0, uint32_t lane) {
  const uint64_t cap_mask = (1ULL << rec.cap_log2) - 1;
  uint32_t rec_start = 0, rec_len = 0, e = 0;
  if (lane < cnt) {
    e = q_e[lane];
    const uint32_t ah = q_ah[lane];
    uint64_t slot = rec.cap_log2 ? (uint64_t)(ah >> (32 - rec.cap_log2)) : 0;
    while (true) {  // open addressing: on to the next slot until the hash or an empty one
      const uint64_t de = rec.dir[slot];
      const uint32_t pi = (uint32_t)de;
      if (pi == kNone) break;
      if ((uint32_t)(de >> 32) == ah) {
        rec_len = rec.post[pi];
        rec_start = pi + 1;
        break;
      }
      slot = (slot + 1) & cap_mask;
    }
  }
  if (!__ballot(rec_len != 0)) return;
  const uint64_t prow = rec_len ? (uint64_t)ent_p[e] * row_words : 0;
  const uint32_t end = (uint32_t)wave_inclusive_scan(rec_len, lane);
  const uint32_t total = (uint32_t)__shfl((int)end, 63, 64);
  ends[lane] = end;
  wave_lds_sync();
  for (uint32_t g0 = 0; g0 < total; g0 += 64) {
    const uint32_t g = g0 + lane < total ? g0 + lane : total - 1;
    uint32_t k = 0;  // the first lane whose postings end after g
#pragma unroll
    for (uint32_t st = 32; st; st >>= 1)
      if (ends[k + st - 1] <= g) k += st;
    const uint32_t first = k ? ends[k - 1] : 0u;
    const uint32_t rs = (uint32_t)__shfl((int)rec_start, (int)k, 64);
    const uint64_t pk = shfl64(prow, k);
    if (g0 + lane < total) {
      const uint32_t pe = rec.post_col[rs + (g - first)];
      const uint32_t col = pe & 0xFFFFFFu;
      if (pe != kNone && (pe >> 24) == colmap && col < groups)
        atomicOr(hits + pk + col_word0 + col / 32, 1u << (col % 32));
    }
  }
  wave_lds_sync();
}

__global__ __launch_bounds__(kRecThreads) void plane_recovery_kernel(
    const uint64_t *ent_h, const uint32_t *ent_p, const uint64_t *n_ptr, uint64_t seed, RecDev rec,
    uint32_t colmap, uint32_t groups, uint32_t *hits, uint32_t row_words, uint32_t col_word0) {
  constexpr int J = 4;  // batches of 64 entries per wave and step, their loads in flight together
  constexpr uint32_t W = kRecThreads / 64;
  __shared__ uint32_t filter[kRecFilterWords];
  __shared__ uint32_t ends_all[W * 64];
  __shared__ uint32_t q_e_all[W * kRecQueue], q_ah_all[W * kRecQueue];
  for (uint32_t i = threadIdx.x; i < kRecFilterWords; i += kRecThreads) filter[i] = rec.filter[i];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t *ends = ends_all + wave * 64;
  uint32_t *q_e = q_e_all + wave * kRecQueue, *q_ah = q_ah_all + wave * kRecQueue;
  const uint64_t n = *n_ptr;  // below 2^32: a chunk's entries
  const uint64_t stride = (uint64_t)gridDim.x * W * 64 * J;
  const uint64_t below = lanemask_lt(lane);
  uint32_t qn = 0;  // queued candidates (wave-uniform)
  for (uint64_t w0 = ((uint64_t)blockIdx.x * W + wave) * 64 * J; w0 < n; w0 += stride) {
    uint64_t hj[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
      const uint64_t e = w0 + (uint64_t)j * 64 + lane;
      hj[j] = e < n ? ent_h[e] : 0;
    }
#pragma unroll
    for (int j = 0; j < J; j++) {
      const uint64_t e = w0 + (uint64_t)j * 64 + lane;
      const uint32_t ah = abandoned_hash(unsplitmix64(hj[j]) - seed, seed);  // mixsplit(key, seed) == h
      bool cand = e < n;
#pragma unroll
      for (uint32_t q = 0; q < kRecFilterProbes; q++) {
        const uint32_t fb = rec_filter_bit(ah, q);
        if (q < rec.filter_probes) cand = cand && ((filter[fb / 32] >> (fb % 32)) & 1u);
      }
      const uint64_t m = __ballot(cand);
      if (cand) {
        const uint32_t at = qn + (uint32_t)__popcll(m & below);
        q_e[at] = (uint32_t)e;
        q_ah[at] = ah;
      }
      qn += (uint32_t)__popcll(m);  // < kRecQueue: at most 63 were queued before
      if (qn >= 64) {
        wave_lds_sync();
        recover_queued(q_e, q_ah, 64, ends, ent_p, rec, colmap, groups, hits, row_words, col_word0,
                       lane);
        // the rest moves to the front of the queue
        uint32_t e2 = 0, ah2 = 0;
        if (64 + lane < qn) {
          e2 = q_e[64 + lane];
          ah2 = q_ah[64 + lane];
        }
        wave_lds_sync();
        if (64 + lane < qn) {
          q_e[lane] = e2;
          q_ah[lane] = ah2;
        }
        qn -= 64;
      }
    }
  }
  if (qn) {
    wave_lds_sync();
    recover_queued(q_e, q_ah, qn, ends, ent_p, rec, colmap, groups, hits, row_words, col_word0, lane);
  }
}

// Bit-sliced counters: plane j holds bit j of each of the 32 counts of a
// lane's word (kPlanes planes: counts up to 1023 between spills).
constexpr int kPlanes = 10;

__device__ __forceinline__ void sum_spill(uint32_t (&pl)[kPlanes], uint32_t (&tot)[32]) {
#pragma clang loop unroll(full)
  for (int b = 0; b < 32; b++) {
    uint32_t c = 0;
#pragma clang loop unroll(full)
    for (int j = 0; j < kPlanes; j++) c |= ((pl[j] >> b) & 1u) << j;
    tot[b] += c;
  }
#pragma clang loop unroll(full)
  for (int j = 0; j < kPlanes; j++) pl[j] = 0;
}

// Adds the counts of a slot (the LP lanes that cover one position's row
// words; lane j of the slot holds word j) to its read (row: counts +
// r * n_colors) or to the colour sums. The slot's LP x 32 per-bit counts go
// through LDS (word-major, `stage`) so that each atomic instruction covers
// LP consecutive bits of a word -- neighbouring colours -- instead of one bit
// of every word; color[i] is the colour of bit i % 32 of word i / 32.
__device__ __forceinline__ void sum_flush(uint32_t (&pl)[kPlanes], uint32_t (&tot)[32],
                                          uint32_t *stage, const uint32_t *color, uint32_t j,
                                          uint32_t lp_log2, uint32_t records, uint32_t *row,
                                          uint64_t *partial) {
  sum_spill(pl, tot);
  uint4 *mine = (uint4 *)(stage + j * 32);
#pragma clang loop unroll(full)
  for (int q = 0; q < 8; q++)
    mine[q] = make_uint4(tot[4 * q], tot[4 * q + 1], tot[4 * q + 2], tot[4 * q + 3]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma clang loop unroll(disable)
  for (uint32_t t = 0; t < 32; t++) {
    const uint32_t i = (t << lp_log2) + j;
    const uint32_t v = stage[i], c = color[i];
    if (v && c != kNone) {
      if (records) atomicAdd(&row[c], v);
      else atomicAdd((unsigned long long *)&partial[c], (unsigned long long)v);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma clang loop unroll(full)
  for (int b = 0; b < 32; b++) tot[b] = 0;
}

// Full and half adders on 32 independent bit positions.
__device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t &s,
                                         uint32_t &carry) {
  s = a ^ b ^ c;
  carry = (a & b) | (c & (a ^ b));
}
__device__ __forceinline__ void half_add(uint32_t a, uint32_t b, uint32_t &s, uint32_t &carry) {
  s = a ^ b;
  carry = a & b;
}

// The per-bit count (0..16) of 16 words as 5 bit planes, by a carry-save
// adder tree: 11 full and 4 half adders.
__device__ __forceinline__ void count16(const uint32_t (&w)[16], uint32_t (&o)[5]) {
  uint32_t s1, c1, s2, c2, s3, c3, s4, c4, s5, c5;
  full_add(w[0], w[1], w[2], s1, c1);
  full_add(w[3], w[4], w[5], s2, c2);
  full_add(w[6], w[7], w[8], s3, c3);
  full_add(w[9], w[10], w[11], s4, c4);
  full_add(w[12], w[13], w[14], s5, c5);
  uint32_t S1, C1, S2, C2, C3;  // ones: s1..s5, w15
  full_add(s1, s2, s3, S1, C1);
  full_add(s4, s5, w[15], S2, C2);
  half_add(S1, S2, o[0], C3);
  uint32_t T1, D1, T2, D2, T3, D3, D4;  // twos: c1..c5, C1, C2, C3
  full_add(c1, c2, c3, T1, D1);
  full_add(c4, c5, C1, T2, D2);
  full_add(C2, C3, T1, T3, D3);
  half_add(T2, T3, o[1], D4);
  uint32_t F1, E1, E2;  // fours: D1..D4
  full_add(D1, D2, D3, F1, E1);
  half_add(F1, D4, o[2], E2);
  half_add(E1, E2, o[3], o[4]);  // eights: E1, E2
}

// pl += x (x: kb planes), rippling through every plane.
template <int KB>
__device__ __forceinline__ void add_planes(uint32_t (&pl)[kPlanes], const uint32_t (&x)[KB]) {
  uint32_t carry = 0;
#pragma clang loop unroll(full)
  for (int j = 0; j < kPlanes; j++) {
    if (j < KB) {
      uint32_t s;
      full_add(pl[j], x[j], carry, s, carry);
      pl[j] = s;
    } else {
      uint32_t s;
      half_add(pl[j], carry, s, carry);
      pl[j] = s;
    }
  }
}

// Per-read sums of the hit rows, in position order. A wave takes a span of
// positions and a group of 64 row words; lane l reads word (l % LP) of
// position (l / LP) + PPI*i (LP lanes per position, PPI = 64 / LP positions
// per step). Every row of the span is read: rows of positions without an
// entry are zero. Blocks of 16 steps in which no lane reaches the end of its
// read are counted by a carry-save adder tree and added to bit-sliced
// counters; other blocks go step by step. Counters spill to per-bit counts
// before they can overflow and are added to the read's counts at its end
// (records) or to partial[colour].
__global__ __launch_bounds__(256) void sweep_sum_kernel(
    const uint32_t *hits, uint32_t row_words, uint32_t lp_log2, const uint32_t *word_color,
    const uint64_t *pos_off, uint64_t n_reads, uint64_t p0, uint64_t n_pos, uint32_t records,
    uint32_t *counts, uint64_t *partial, uint32_t n_colors) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t groups = (row_words + 63) / 64;
  const uint64_t spans = (n_pos + kSumSpan - 1) / kSumSpan;
  // a block's four waves share one group of row words (its colour table is
  // staged in LDS) and take four consecutive spans; wave-uniform (scalar)
  const uint32_t g = blockIdx.x % groups;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t sp = (uint64_t)(blockIdx.x / groups) * 4 + wave;
  __shared__ uint32_t color_lds[64 * 32];
  __shared__ uint32_t flush_lds[4][64 * 32];  // per wave: its slots' per-bit counts
  for (uint32_t i = threadIdx.x; i < 64 * 32; i += 256) {
    const uint32_t wi = g * 64 + i / 32;
    color_lds[i] = wi < row_words ? word_color[(uint64_t)g * 64 * 32 + i] : kNone;
  }
  __syncthreads();
  if (sp >= spans) return;
  const uint32_t LP = 1u << lp_log2, PPI = 64 / LP;
  const uint32_t slot = lane >> lp_log2, j = lane & (LP - 1);
  const uint32_t w = g * 64 + j;
  const bool live = w < row_words;
  uint32_t *stage = flush_lds[wave] + slot * LP * 32;
  const uint64_t q0 = sp * kSumSpan, q1 = q0 + kSumSpan < n_pos ? q0 + kSumSpan : n_pos;
  uint64_t first = q0 + slot < q1 ? q0 + slot : q1 - 1;
  uint64_t r = 0;
  if (records) {  // the read of the span's first position, then of the slot's
    r = wave_find_segment(pos_off, n_reads, p0 + q0, lane);
    while (pos_off[r + 1] <= p0 + first) r++;
  }
  uint64_t next = records ? pos_off[r + 1] - p0 : ~0ULL;  // relative end of read r
  uint32_t pl[kPlanes], tot[32];
#pragma clang loop unroll(full)
  for (int k = 0; k < kPlanes; k++) pl[k] = 0;
#pragma clang loop unroll(full)
  for (int b = 0; b < 32; b++) tot[b] = 0;
  uint32_t fill = 0;
  constexpr int U = 16;  // steps whose loads are in flight together
  const uint64_t stride = (uint64_t)PPI * row_words;  // words between a lane's steps
  for (uint64_t pp = q0; pp < q1; pp += (uint64_t)PPI * U) {
    if (fill > (1u << kPlanes) - 1 - U) {
      sum_spill(pl, tot);
      fill = 0;
    }
    fill += U;
    const uint64_t last = pp + slot + (uint64_t)(U - 1) * PPI;
    const bool cross = records && (last < q1 ? last : q1 - 1) >= next;
    if (!__ballot(cross)) {
      // no lane reaches the end of its read in these U steps
      uint32_t xw[U];
      const uint32_t *at = hits + (pp + slot) * row_words + w;
#pragma clang loop unroll(full)
      for (int u = 0; u < U; u++) {
        const uint64_t p = pp + slot + (uint64_t)u * PPI;
        xw[u] = live && p < q1 ? at[u * stride] : 0u;
      }
      uint32_t c5[5];
      count16(xw, c5);
      add_planes<5>(pl, c5);
    } else {
#pragma clang loop unroll(disable)
      for (int u = 0; u < U; u++) {
        const uint64_t p = pp + slot + (uint64_t)u * PPI;
        uint32_t x = 0;
        if (p < q1) {
          if (p >= next) {
            sum_flush(pl, tot, stage, color_lds, j, lp_log2, records,
                      counts + r * n_colors, partial);
            while (pos_off[r + 1] - p0 <= p) r++;
            next = pos_off[r + 1] - p0;
          }
          if (live) x = hits[p * row_words + w];
        }
        const uint32_t x1[1] = {x};
        add_planes<1>(pl, x1);
      }
    }
  }
  sum_flush(pl, tot, stage, color_lds, j, lp_log2, records, counts + r * n_colors, partial);
}

uint32_t bits_for(uint64_t v) {  // bits to hold values 0..v
  uint32_t b = 1;
  while (b < 64 && (v >> b)) b++;
  return b;
}

}  // namespace

static uint32_t slice_bytes_for(const BankDev &bk) {
  return (uint32_t)std::min<uint64_t>(kSliceBytes, (bk.plane_bytes + 63) / 64 * 64);
}

// Row-slice bytes of the bit-plane lookup for a bank of b-bit fields, 0 when
// it does not apply: banks without merged groups, H in {2, 3, 4, 8}, the
// window of H x L slices within the kernel's LDS window. The largest slice
// that fits: at b in {4, 8, 16} 512 bytes (32 pieces) for H <= 4 and 256 for
// H = 8; at b = 32, 256 and 128; at b <= 2, 128 bytes (a 1024-column row at
// b = 1).
static uint32_t plane_slice_bytes(const BankDev &bk, uint32_t B) {
  if (bk.first_merged != bk.groups || bk.groups == 0) return 0;
  if (bk.hashes != 2 && bk.hashes != 3 && bk.hashes != 4 && bk.hashes != 8) return 0;
  uint32_t window = 0, sb = 0;
  switch (B) {
    case 1: case 2: window = plane_window(1); sb = 128; break;
    case 4: case 8: case 16: window = plane_window(8); sb = bk.hashes <= 4 ? 512 : 256; break;
    case 32: window = plane_window(32); sb = bk.hashes <= 4 ? 256 : 128; break;
    default: return 0;
  }
  if ((uint64_t)bk.hashes * bk.geo.length * sb > window) return 0;
  return sb;
}

static bool planes_enabled() {
  const char *e = getenv("ZG_SWEEP_PLANES");  // 0: always the staged lookup
  return !e || strcmp(e, "0") != 0;
}

// Whether the staged lookup's window (H x L rows of a 512-byte slice) fits LDS.
static bool stage_fits(const BankDev &bk) {
  return (uint64_t)bk.hashes * bk.geo.length * slice_bytes_for(bk) <= kStageWindow;
}

// Indexes the sweep serves: every seed mode but findere, banks resident, and
// every bank's segment window within LDS for the staged lookup
// (H * L * slice <= 128 KiB, L <= 256, H <= 8) or for the bit-plane lookup
// (then the bank must get its bit-plane copy).
bool sweep_layout_ok(const zg_index *idx) {
  if (idx->m.index_mode == 3 || idx->dbanks.empty()) return false;
  const uint32_t B = (uint32_t)idx->m.fingerprint_bits;
  for (const DeviceBank &db : idx->dbanks) {
    const BankDev &bk = db.view;
    if (bk.geo.length > 256 || bk.hashes > 8 || bk.geo.count < 1 || bk.groups == 0) return false;
    if (!stage_fits(bk) && !(planes_enabled() && plane_slice_bytes(bk, B))) return false;
  }
  return true;
}

// Bank rows any lookup reads: (count + H - 1) segments of L rows.
static uint64_t lookup_rows(const BankDev &bk) {
  return (bk.geo.count + bk.hashes - 1) * bk.geo.length;
}

template <int B>
static void launch_bit_planes(const BankDev &bk, uint64_t rows, uint64_t pitch, uint8_t *out) {
  const uint32_t g32 = (bk.groups + 31) / 32;
  const uint64_t n = rows * g32;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(65536, (n + 255) / 256);
  bit_planes_kernel<B><<<blocks, 256, 0, 0>>>(bk.data, bk.row_bytes, rows, bk.groups, g32, pitch, out);
}

// The bit-plane copy of a bank: slices of sb bytes, n_slices per row.
static void build_planes(DeviceBank &db, uint32_t B, uint32_t sb, uint32_t n_slices) {
  const BankDev &bk = db.view;
  const uint64_t rows = lookup_rows(bk), pitch = (uint64_t)n_slices * sb;
  db.planes.alloc(rows * pitch);
  check_hip(hipMemsetAsync(db.planes.p, 0, rows * pitch, 0), "bit planes");
  uint8_t *out = db.planes.as<uint8_t>();
  switch (B) {
    case 1: launch_bit_planes<1>(bk, rows, pitch, out); break;
    case 2: launch_bit_planes<2>(bk, rows, pitch, out); break;
    case 4: launch_bit_planes<4>(bk, rows, pitch, out); break;
    case 8: launch_bit_planes<8>(bk, rows, pitch, out); break;
    case 16: launch_bit_planes<16>(bk, rows, pitch, out); break;
    default: launch_bit_planes<32>(bk, rows, pitch, out); break;
  }
  check_hip(hipGetLastError(), "bit planes");
  check_hip(hipStreamSynchronize(0), "bit planes");
  db.planes_pitch = pitch;
}

template <int B>
static void launch_planes(const StageArgs &a, uint32_t cus, hipStream_t s) {
  // one 1024-thread block per CU (wide kernels) or two 512-thread blocks
  const uint32_t blocks = std::max<uint32_t>(a.xcds, plane_wide(B) ? cus : 2 * cus);
  constexpr uint32_t T = plane_threads(B);
  constexpr int NP4 = plane_wide(B) ? 32 : (B == 32 ? 16 : 8);  // pieces at H <= 4
  constexpr int NP8 = B <= 2 ? 8 : NP4 / 2;                      // pieces at H = 8
  ensure(a.slice_bytes == (a.bank.hashes <= 4 ? NP4 : NP8) * 16u, "bit-plane lookup: slice size",
         ZG_ERR_INVALID);
  switch (a.bank.hashes) {
    case 2: plane_lookup_kernel<B, 2, NP4><<<blocks, T, 0, s>>>(a); break;
    case 3: plane_lookup_kernel<B, 3, NP4><<<blocks, T, 0, s>>>(a); break;
    case 4: plane_lookup_kernel<B, 4, NP4><<<blocks, T, 0, s>>>(a); break;
    case 8: plane_lookup_kernel<B, 8, NP8><<<blocks, T, 0, s>>>(a); break;
    default: fail(ZG_ERR_INVALID, "bit-plane lookup: hash count");
  }
}

template <int B>
static void launch_stage(const StageArgs &a, uint32_t blocks, hipStream_t s) {
  switch (a.bank.hashes) {
    case 2: stage_lookup_kernel<B, 2><<<blocks, kStageThreads, 0, s>>>(a); break;
    case 3: stage_lookup_kernel<B, 3><<<blocks, kStageThreads, 0, s>>>(a); break;
    case 4: stage_lookup_kernel<B, 4><<<blocks, kStageThreads, 0, s>>>(a); break;
    default: stage_lookup_kernel<B, 0><<<blocks, kStageThreads, 0, s>>>(a); break;
  }
}

static std::vector<uint32_t> fetch_u32(const uint32_t *p, uint64_t n) {
  std::vector<uint32_t> v(std::max<uint64_t>(1, n));
  if (n) check_hip(hipMemcpy(v.data(), p, n * 4, hipMemcpyDeviceToHost), "sweep tables");
  return v;
}

// The per-index tables of the sweep (built on first use, outside any
// capture): size classes, the hit-row layout and the colour of every row bit.
// Row layout: each class's column region (bit-plane classes: n_slices x SG
// words, column c at bit c; others: n_slices x 512/b bytes, column c at the
// bit of its lane field), then every class's member bits.
static void ensure_sweep_tables(zg_index *idx) {
  const int allow = planes_enabled() ? 1 : 0;
  std::lock_guard<std::mutex> lock(idx->sw_mu);  // server slots share the index
  if (idx->sw_row_words && idx->sw_planes == allow) return;
  for (DeviceBank &db : idx->dbanks) {  // tables of the other setting
    db.planes.reset();
    db.planes_pitch = 0;
  }
  idx->sw_planes = allow;
  const uint32_t B = (uint32_t)idx->m.fingerprint_bits;
  std::vector<SweepClass> classes;
  for (uint32_t bi = 0; bi < idx->dbanks.size(); bi++) {
    const BankDev &bk = idx->dbanks[bi].view;
    bool placed = false;
    for (SweepClass &c : classes) {
      const BankDev &o = idx->dbanks[c.banks[0]].view;
      if (o.colmap == bk.colmap && o.geo.count == bk.geo.count && o.geo.length == bk.geo.length &&
          o.row_bytes == bk.row_bytes && o.groups == bk.groups) {
        c.banks.push_back(bi);
        placed = true;
        break;
      }
    }
    if (!placed) {
      SweepClass c;
      c.banks.push_back(bi);
      c.count = bk.geo.count;
      c.slice_bytes = slice_bytes_for(bk);
      c.n_slices = (uint32_t)((bk.plane_bytes + c.slice_bytes - 1) / c.slice_bytes);
      classes.push_back(c);
    }
  }
  // bit-plane copies for the classes the plane lookup serves, while device
  // memory allows (a copy costs about one bank; a quarter of the free memory
  // stays for the chunk buffers)
  for (SweepClass &c : classes) {
    const BankDev &bk = idx->dbanks[c.banks[0]].view;
    const uint32_t sb = allow ? plane_slice_bytes(bk, B) : 0;
    if (!sb) continue;
    const uint32_t sg = sb / (4 * B), g32 = (bk.groups + 31) / 32;
    const uint32_t n_slices = (g32 + sg - 1) / sg;
    const uint64_t need = lookup_rows(bk) * n_slices * (uint64_t)sb * c.banks.size();
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      free_b = 0;
    }
    // a quarter of the free memory stays for the chunks, unless only the bit
    // planes can serve the class (the staged window does not fit LDS)
    const bool only = !stage_fits(bk);
    if (need > (only ? free_b : free_b / 4 * 3)) {
      ensure(!only, "the sorted sweep needs a bit-plane copy of the banks that does not fit the "
                    "free device memory", ZG_ERR_NOMEM);
      continue;
    }
    for (uint32_t bi : c.banks) build_planes(idx->dbanks[bi], B, sb, n_slices);
    c.planes = true;
    c.slice_bytes = sb;
    c.n_slices = n_slices;
  }
  uint32_t col_words = 0, members = 0;
  for (SweepClass &c : classes) {
    const BankDev &bk = idx->dbanks[c.banks[0]].view;
    c.col_word0 = col_words;
    if (c.planes)  // 32 columns per word, regions in whole 16-byte pieces
      col_words += (c.n_slices * (c.slice_bytes / (4 * B)) + 3) / 4 * 4;
    else
      col_words += c.n_slices * (kSliceBytes / B) / 4;  // 64 lanes x 64/b bits per slice
    c.members = bk.groups > bk.first_merged
                    ? fetch_u32(bk.mg_off, bk.groups - bk.first_merged + 1)[bk.groups - bk.first_merged]
                    : 0;
    c.member_bit0 = members;  // relative to the member region until the columns are placed
    members += c.members;
  }
  const uint32_t row_words = col_words + (members + 31) / 32;
  std::vector<uint32_t> color(row_words * 32ULL, kNone);
  for (SweepClass &c : classes) {
    const BankDev &bk = idx->dbanks[c.banks[0]].view;
    c.member_bit0 += col_words * 32;
    const std::vector<uint32_t> col_color = fetch_u32(bk.col_color, bk.groups);
    // column col = bit col of the region (bit planes), or lane fields: bit
    // (col / F) * F + field_bit(col % F)
    const uint32_t F = 64 / B;
    for (uint32_t col = 0; col < bk.first_merged; col++) {
      const uint32_t f = col % F;
      const uint32_t bit = c.planes ? col : col - f + (B == 8 ? field_bit<8>(f) : f);
      color[c.col_word0 * 32ULL + bit] = col_color[col];
    }
    const std::vector<uint32_t> mem = fetch_u32(bk.mg_members, c.members);
    for (uint32_t i = 0; i < c.members; i++) color[c.member_bit0 + (uint64_t)i] = mem[i];
  }
  idx->sw_bit_color.alloc(color.size() * 4);
  check_hip(hipMemcpy(idx->sw_bit_color.p, color.data(), color.size() * 4, hipMemcpyHostToDevice),
            "sweep tables");
  idx->sw_classes = classes;
  idx->sw_col_words = col_words;
  idx->sw_members = members;
  idx->sw_row_words = row_words;
}

// The sort of one class: LSD passes on the first segment, then the segment
// starts. Returns the buffer pair holding the sorted entries; *n_dev is the
// entry count in device memory.
template <class Src>
static int sort_class(zg_batch *b, Src first_pass, uint64_t count, uint64_t n_items,
                      uint64_t *listed, const uint64_t **n_dev, hipStream_t s) {
  const uint32_t seg_bits = bits_for(count - 1);
  const uint32_t passes = (seg_bits + kMaxDigitBits - 1) / kMaxDigitBits;
  const uint32_t dbits = (seg_bits + passes - 1) / passes, nb = 1u << dbits;
  const uint64_t tiles = std::max<uint64_t>(1, (n_items + kSortTile - 1) / kSortTile);
  const uint64_t hn = (uint64_t)nb * tiles + 1;
  const size_t tb = scan_temp_bytes_u32_to_u64(hn);
  const uint64_t *n_ptr = b->sw_off.as<uint64_t>() + (hn - 1);  // the scan's total
  *n_dev = n_ptr;
  int cur = 0;
  for (uint32_t pass = 0; pass < passes; pass++) {
    const uint32_t shift = pass * dbits;
    EntSource es{b->sw_h[cur].as<uint64_t>(), b->sw_p[cur].as<uint32_t>(), n_ptr, 0};
    if (pass == 0)
      radix_hist_kernel<Src><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
          first_pass, count, shift, dbits, tiles, b->sw_hist.as<uint32_t>(), listed);
    else
      radix_hist_kernel<EntSource><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
          es, count, shift, dbits, tiles, b->sw_hist.as<uint32_t>(), nullptr);
    launch_exclusive_scan_u32_to_u64(b->sw_hist.as<uint32_t>(), b->sw_off.as<uint64_t>(), hn,
                                     b->scan_temp.p, tb, s);
    // the last pass also marks each segment's first entry
    uint32_t *bounds = nullptr;
    if (pass + 1 == passes) {
      bounds = b->sw_start.as<uint32_t>();
      bounds_init_kernel<<<(uint32_t)std::min<uint64_t>(1024, count / 256 + 1), 256, 0, s>>>(
          bounds, count, n_ptr);
    }
    if (pass == 0) {
      radix_scatter_kernel<Src><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
          first_pass, count, shift, dbits, tiles, b->sw_off.as<uint64_t>(),
          b->sw_h[0].as<uint64_t>(), b->sw_p[0].as<uint32_t>(), bounds);
      cur = 0;
    } else {
      radix_scatter_kernel<EntSource><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
          es, count, shift, dbits, tiles, b->sw_off.as<uint64_t>(), b->sw_h[1 - cur].as<uint64_t>(),
          b->sw_p[1 - cur].as<uint32_t>(), bounds);
      cur = 1 - cur;
    }
  }
  bounds_fill_kernel<<<1, 1024, 0, s>>>(b->sw_start.as<uint32_t>(), count);
  return cur;
}

// Device bytes of each buffer of a chunk of `most` items (entries twice, the
// listed-position bitmap, tile histograms and offsets, scan space, segment
// starts, work counters, hit rows), reserved in that order.
static std::vector<uint64_t> chunk_buffer_bytes(const zg_index *idx, uint64_t most, bool keys) {
  uint64_t max_count = 1, max_hist = 1;
  for (const SweepClass &c : idx->sw_classes) {
    max_count = std::max(max_count, c.count);
    const uint32_t seg_bits = bits_for(c.count - 1);
    const uint32_t passes = (seg_bits + kMaxDigitBits - 1) / kMaxDigitBits;
    const uint32_t dbits = (seg_bits + passes - 1) / passes;
    max_hist = std::max<uint64_t>(
        max_hist, ((uint64_t)1 << dbits) * std::max<uint64_t>(1, (most + kSortTile - 1) / kSortTile) + 1);
  }
  return {most * 8 + 64, most * 8 + 64, most * 4 + 64, most * 4 + 64,
          keys ? 0 : ((most + 63) / 64 + 1) * 8, max_hist * 4, max_hist * 8,
          scan_temp_bytes_u32_to_u64(max_hist), (max_count + 1) * 4, 32 * kXcds * 4 + 64,
          std::max<uint64_t>(1, most * idx->sw_row_words) * 4};
}

static void reserve_chunk(zg_index *idx, zg_batch *b, uint64_t most, bool keys) {
  const std::vector<uint64_t> n = chunk_buffer_bytes(idx, most, keys);
  DevBuf *bufs[] = {&b->sw_h[0], &b->sw_h[1], &b->sw_p[0], &b->sw_p[1], &b->sw_listed,
                    &b->sw_hist, &b->sw_off, &b->scan_temp, &b->sw_start, &b->sw_queue,
                    &b->fd_hits};
  for (size_t i = 0; i < n.size(); i++)
    if (n[i]) bufs[i]->reserve(n[i]);
}

void sweep_prepare(zg_index *idx) { ensure_sweep_tables(idx); }

uint64_t sweep_chunk_bytes(zg_index *idx, uint64_t positions) {
  ensure_sweep_tables(idx);
  uint64_t n = 0;
  for (uint64_t v : chunk_buffer_bytes(idx, positions, false)) n += v;
  return n;
}

void sweep_reserve(zg_index *idx, zg_batch *b, uint64_t positions) {
  ensure_sweep_tables(idx);
  reserve_chunk(idx, b, positions, false);
}

uint32_t sweep_row_words(zg_index *idx) {
  ensure_sweep_tables(idx);
  return idx->sw_row_words;
}

// The lookups of a batch as the sweep, every chunk enqueued on b->stream with
// no host synchronization (sizes come from the host's read offsets). Entries
// are the valid windows (positions) or the explicit keys of a chunk of whole
// reads; per chunk, events bracket each class's sort and lookups and the sum.
void sweep_run(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  const Manifest &m = idx->m;
  ensure_sweep_tables(idx);
  const uint32_t row_words = idx->sw_row_words;
  const uint32_t B = (uint32_t)m.fingerprint_bits;
  const bool keys = b->explicit_keys;
  const bool strobes = m.has_randstrobemers;
  // rows of counts: reads, or both orientations of each read for randstrobes
  const uint64_t n_rows = strobes ? 2 * b->n_reads : b->n_reads;
  // chunk size: hit rows of at most 160 GiB, entries below 2^32, and the
  // chunk's buffers (hit row + 24 B of entries per item) within 3/4 of the
  // device memory free or already held by them
  uint64_t hit_budget = kSweepHitBytes;
  if (const char *e = getenv("ZG_SWEEP_HIT_GIB"))  // hit rows of one chunk, GiB
    hit_budget = std::max<uint64_t>(1, strtoull(e, nullptr, 10)) << 30;
  if (b->sw_hit_budget) hit_budget = std::min(hit_budget, b->sw_hit_budget);
  uint64_t cap = std::min<uint64_t>((1ULL << 32) - 64, hit_budget / (row_words * 4ULL));
  size_t free_b = 0, total_b = 0;
  if (b->sw_capturing) {
    cap = b->sw_cap;  // the eager run's chunks: every buffer already holds them
  } else if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    uint64_t held = 0;
    for (const DevBuf *d : {&b->fd_hits, &b->sw_h[0], &b->sw_h[1], &b->sw_p[0], &b->sw_p[1],
                            &b->sw_listed, &b->sw_hist, &b->sw_off, &b->scan_temp})
      held += d->bytes;
    cap = std::min<uint64_t>(cap, std::max<uint64_t>(1 << 16, (free_b + held) / 4 * 3 /
                                                                   (row_words * 4ULL + 32)));
  } else {
    (void)hipGetLastError();
  }
  if (const char *e = getenv("ZG_SWEEP_CHUNK_POSITIONS"))  // smaller chunks
    cap = std::min<uint64_t>(cap, std::max<uint64_t>(64, strtoull(e, nullptr, 10)));
  // hit rows already held cover this many positions: a cap less than 1/8
  // above it is not worth reallocating the chunk buffers for (free memory
  // as reported moves by a few hundred MB between runs)
  const uint64_t held_cap = b->fd_hits.bytes / (row_words * 4ULL);
  if (!b->sw_capturing && held_cap && cap > held_cap && cap - held_cap < held_cap / 8)
    cap = held_cap;
  b->sw_cap = cap;
  const std::vector<uint64_t> &off = b->h_seg_off;  // row starts in the item space
  std::vector<uint64_t> cuts{0};
  uint64_t most = 0;
  while (cuts.back() < n_rows) {
    const uint64_t r0 = cuts.back();
    uint64_t r1 = r0 + 1;
    while (r1 < n_rows && off[r1 + 1] - off[r0] <= cap) r1++;
    ensure(off[r1] - off[r0] < (1ULL << 32) - 64,
           "a read is too long for the sorted sweep (2^32 seeds per chunk)", ZG_ERR_INVALID);
    most = std::max(most, off[r1] - off[r0]);
    cuts.push_back(r1);
  }
  reserve_chunk(idx, b, most, keys);
  const uint32_t n_chunks = (uint32_t)cuts.size() - 1;
  const size_t ev_per_chunk = 2 + 3 * idx->sw_classes.size();
  while (b->sw_ev.size() < ev_per_chunk * n_chunks) {
    hipEvent_t e;
    check_hip(hipEventCreate(&e), "hipEventCreate");
    b->sw_ev.push_back(e);
  }
  unsigned long long *folded = (unsigned long long *)(b->sw_queue.as<uint32_t>() + 32 * kXcds);
  launch_fill_u64((uint64_t *)folded, 1, 0, s);
  StageArgs sa{};
  sa.rec = idx->rec_view;
  if (b->flags & ZG_QUERY_NO_RECOVERY) sa.rec.enabled = 0;
  sa.seed = m.seed;
  sa.row_words = row_words;
  sa.hits = b->fd_hits.as<uint32_t>();
  sa.start = b->sw_start.as<uint32_t>();
  sa.queue = b->sw_queue.as<uint32_t>();
  sa.xcds = idx->cu_count >= kXcds ? kXcds : 1;
  sa.folded = folded;
  const uint32_t stage_blocks = std::max<uint32_t>(sa.xcds, idx->cu_count);
  // entry ranges per unit of the bit-plane lookup: fewer segments in flight
  // per XCD, so more hit rows are whole in L2 before they are written back
  // hit rows of the bit-plane lookup through LDS, whole per store (1, the
  // default) or stored 16 bytes per lane (0)
  sa.row_stage = 1;
  if (const char *e = getenv("ZG_SWEEP_ROW_STAGE")) sa.row_stage = atoi(e) ? 1 : 0;
  uint32_t plane_parts = 1;
  if (const char *e = getenv("ZG_SWEEP_PARTS"))
    plane_parts = (uint32_t)std::min<unsigned long>(64, std::max<unsigned long>(1, strtoul(e, nullptr, 10)));
  // sum: lanes per item, the row's words rounded up to a power of two (at most 64)
  uint32_t lp_bits = 0;
  while ((1u << lp_bits) < std::min<uint32_t>(64, row_words)) lp_bits++;
  b->sweep_chunks = n_chunks;
  b->sweep_classes = (uint32_t)idx->sw_classes.size();
  for (uint32_t ci = 0; ci < n_chunks; ci++) {
    const uint64_t I0 = off[cuts[ci]], n_items = off[cuts[ci + 1]] - I0;
    hipEvent_t *ev = &b->sw_ev[ev_per_chunk * ci];
    check_hip(hipEventRecord(ev[0], s), "hipEventRecord");
    if (idx->sw_members)
      zero_members_kernel<<<16384, 256, 0, s>>>(b->fd_hits.as<uint32_t>(), n_items, row_words,
                                                 idx->sw_col_words, row_words - idx->sw_col_words);
    for (size_t k = 0; k < idx->sw_classes.size(); k++) {
      const SweepClass &c = idx->sw_classes[k];
      int cur;
      const uint64_t *n_dev = nullptr;  // the class's entry count
      if (keys) {
        KeySource ks{b->keys.as<uint64_t>(), I0, n_items, m.seed};
        cur = sort_class(b, ks, c.count, n_items, nullptr, &n_dev, s);
      } else {
        PosSource ps{b->codes.as<uint64_t>(), b->invalid.as<uint64_t>(), I0, n_items, (uint32_t)m.k,
                     m.seed, m.index_mode == 2 ? m.modimizer_sampling : 0, idx->spaced};
        cur = sort_class(b, ps, c.count, n_items, k == 0 ? b->sw_listed.as<uint64_t>() : nullptr,
                         &n_dev, s);
      }
      check_hip(hipEventRecord(ev[1 + 3 * k], s), "hipEventRecord");
      sa.ent_h = b->sw_h[cur].as<uint64_t>();
      sa.ent_p = b->sw_p[cur].as<uint32_t>();
      b->sw_last_p = sa.ent_p;
      b->sw_last_n = n_dev;
      b->sw_last_items = n_items;
      sa.slice_bytes = c.slice_bytes;
      sa.n_slices = c.n_slices;
      sa.col_word0 = c.col_word0;
      sa.member_bit0 = c.member_bit0;
      for (uint32_t bi : c.banks) {
        sa.bank = idx->dbanks[bi].view;
        launch_fill_u32(sa.queue, 32 * kXcds, 0, s);
        if (c.planes) {
          sa.planes = idx->dbanks[bi].planes.as<uint8_t>();
          sa.planes_pitch = idx->dbanks[bi].planes_pitch;
          sa.parts = plane_parts;
          switch (B) {
            case 1: launch_planes<1>(sa, idx->cu_count, s); break;
            case 2: launch_planes<2>(sa, idx->cu_count, s); break;
            case 4: launch_planes<4>(sa, idx->cu_count, s); break;
            case 8: launch_planes<8>(sa, idx->cu_count, s); break;
            case 16: launch_planes<16>(sa, idx->cu_count, s); break;
            default: launch_planes<32>(sa, idx->cu_count, s); break;
          }
          continue;
        }
        sa.planes = nullptr;
        switch (B) {
          case 1: launch_stage<1>(sa, stage_blocks, s); break;
          case 2: launch_stage<2>(sa, stage_blocks, s); break;
          case 4: launch_stage<4>(sa, stage_blocks, s); break;
          case 8: launch_stage<8>(sa, stage_blocks, s); break;
          case 16: launch_stage<16>(sa, stage_blocks, s); break;
          default: launch_stage<32>(sa, stage_blocks, s); break;
        }
      }
      check_hip(hipEventRecord(ev[2 + 3 * k], s), "hipEventRecord");
      if (c.planes && sa.rec.enabled) {
        const BankDev &bk0 = idx->dbanks[c.banks[0]].view;
        plane_recovery_kernel<<<(uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(idx->cu_count, n_items / 4096 + 1)),
                                kRecThreads, 0, s>>>(
            sa.ent_h, sa.ent_p, n_dev, sa.seed, sa.rec, bk0.colmap, bk0.groups, sa.hits, row_words,
            c.col_word0);
      }
      check_hip(hipEventRecord(ev[3 + 3 * k], s), "hipEventRecord");
    }
    const uint64_t groups = (row_words + 63) / 64;
    const uint64_t blocks = groups * (((n_items + kSumSpan - 1) / kSumSpan + 3) / 4);
    if (n_items && !keys)
      zero_unlisted_kernel<<<(uint32_t)std::min<uint64_t>(16384, ((n_items + 63) / 64 + 3) / 4), 256,
                             0, s>>>(b->fd_hits.as<uint32_t>(), row_words,
                                     b->sw_listed.as<uint64_t>(), n_items);
    if (n_items)
      sweep_sum_kernel<<<(uint32_t)blocks, 256, 0, s>>>(
          b->fd_hits.as<uint32_t>(), row_words, lp_bits, idx->sw_bit_color.as<uint32_t>(),
          keys ? b->key_off.as<uint64_t>() : b->pos_off.as<uint64_t>(), n_rows, I0, n_items,
          (b->records || strobes) ? 1 : 0, b->counts.as<uint32_t>(), b->partial.as<uint64_t>(),
          (uint32_t)m.colors.size());
    check_hip(hipEventRecord(ev[ev_per_chunk - 1], s), "hipEventRecord");
  }
  if (strobes) launch_strand_max(b->counts.as<uint32_t>(), b->n_reads * m.colors.size(), s);
}

// The lookups' hit-row stores alone, for the n entries in the sweep's order
// (segment, then position): the row of position p[i] (row_words words)
// written whole by 2^lpr_log2 neighbouring lanes of one store instruction,
// 16 bytes per lane -- the store ceiling of the lookups' position pattern.
__global__ __launch_bounds__(256) void store_probe_kernel(const uint32_t *p, const uint64_t *n_ptr,
                                                          uint32_t *hits, uint32_t row_words,
                                                          uint32_t lpr_log2) {
  const uint64_t n = *n_ptr;
  const uint32_t pieces = row_words / 4, lpr = 1u << lpr_log2;
  const uint32_t j = threadIdx.x & (lpr - 1);
  const uint64_t groups = ((uint64_t)gridDim.x * blockDim.x) >> lpr_log2;
  for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> lpr_log2; i < n;
       i += groups) {
    uint4 *row = (uint4 *)(hits + (uint64_t)p[i] * row_words);
    for (uint32_t k = j; k < pieces; k += lpr) row[k] = make_uint4((uint32_t)i, k, 0, 0);
  }
}

double sweep_store_probe(zg_batch *b) {
  zg_index *idx = b->idx;
  ensure(b->sw_last_p && b->sw_last_n && idx->sw_row_words % 4 == 0,
         "store probe: no sorted sweep has run on this batch", ZG_ERR_INVALID);
  hipStream_t s = b->stream;
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "hipEventCreate");
  check_hip(hipEventCreate(&e1), "hipEventCreate");
  // lanes per row: the row's 16-byte pieces rounded up to a power of two, at most 64
  uint32_t lpr_log2 = 0;
  while ((1u << lpr_log2) < std::min<uint32_t>(64, idx->sw_row_words / 4)) lpr_log2++;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(
      idx->cu_count * 16, (b->sw_last_items << lpr_log2) / 256 + 1);
  store_probe_kernel<<<blocks, 256, 0, s>>>(b->sw_last_p, b->sw_last_n, b->fd_hits.as<uint32_t>(),
                                            idx->sw_row_words, lpr_log2);  // warm-up
  check_hip(hipEventRecord(e0, s), "hipEventRecord");
  store_probe_kernel<<<blocks, 256, 0, s>>>(b->sw_last_p, b->sw_last_n, b->fd_hits.as<uint32_t>(),
                                            idx->sw_row_words, lpr_log2);
  check_hip(hipEventRecord(e1, s), "hipEventRecord");
  check_hip(hipStreamSynchronize(s), "store probe");
  uint64_t n = 0;
  check_hip(hipMemcpy(&n, b->sw_last_n, 8, hipMemcpyDeviceToHost), "store probe");
  const double ms = elapsed_ms(e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)n * idx->sw_row_words * 4 / (ms / 1000) / 1e9;
}

// Phase times of the last sweep (its events have completed). Per chunk:
// start, then per class the ends of its sort, its lookups and its recovery
// pass, then the end of the sums.
void sweep_times(zg_batch *b, double &sort_ms, double &lookup_ms, double &recovery_ms,
                 double &sum_ms) {
  sort_ms = lookup_ms = recovery_ms = sum_ms = 0;
  const size_t per = 2 + 3 * (size_t)b->sweep_classes;
  for (uint64_t ci = 0; ci < b->sweep_chunks; ci++) {
    hipEvent_t *ev = &b->sw_ev[per * ci];
    for (uint32_t k = 0; k < b->sweep_classes; k++) {
      sort_ms += elapsed_ms(ev[3 * k], ev[1 + 3 * k]);
      lookup_ms += elapsed_ms(ev[1 + 3 * k], ev[2 + 3 * k]);
      recovery_ms += elapsed_ms(ev[2 + 3 * k], ev[3 + 3 * k]);
    }
    sum_ms += elapsed_ms(ev[per - 2], ev[per - 1]);
  }
}

uint64_t sweep_folded(zg_batch *b) {
  unsigned long long v = 0;
  if (b->sw_queue.p)
    check_hip(hipMemcpy(&v, b->sw_queue.as<uint32_t>() + 32 * kXcds, 8, hipMemcpyDeviceToHost),
              "hipMemcpy");
  return v;
}

// Building blocks run alone, for the GPU check against their CPU
// restatement (tests/sweep_pieces.cpp). Entries (h[i], p[i]), i < *n_dev,
// sorted stably by their first segment mulhi(h, count); returns the buffer
// pair holding the result (0: h1/p1, 1: h2/p2).
int sweep_sort_entries(const uint64_t *h, const uint32_t *p, const uint64_t *n_dev, uint64_t n_max,
                       uint64_t count, uint64_t *h1, uint32_t *p1, uint64_t *h2, uint32_t *p2,
                       hipStream_t s) {
  const uint32_t seg_bits = bits_for(count - 1);
  const uint32_t passes = (seg_bits + kMaxDigitBits - 1) / kMaxDigitBits;
  const uint32_t dbits = (seg_bits + passes - 1) / passes, nb = 1u << dbits;
  const uint64_t tiles = std::max<uint64_t>(1, (n_max + kSortTile - 1) / kSortTile);
  const uint64_t hn = (uint64_t)nb * tiles + 1;
  DevBuf hist, off, temp, nbuf;
  hist.alloc(hn * 4);
  off.alloc(hn * 8);
  const size_t tb = scan_temp_bytes_u32_to_u64(hn);
  temp.alloc(tb);
  uint64_t *hs[2] = {h1, h2};
  uint32_t *ps[2] = {p1, p2};
  EntSource es{h, p, n_dev, 0};
  int cur = -1;
  for (uint32_t pass = 0; pass < passes; pass++) {
    radix_hist_kernel<EntSource><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
        es, count, pass * dbits, dbits, tiles, hist.as<uint32_t>(), nullptr);
    launch_exclusive_scan_u32_to_u64(hist.as<uint32_t>(), off.as<uint64_t>(), hn, temp.p, tb, s);
    const int nxt = cur < 0 ? 0 : 1 - cur;
    radix_scatter_kernel<EntSource><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
        es, count, pass * dbits, dbits, tiles, off.as<uint64_t>(), hs[nxt], ps[nxt], nullptr);
    es = EntSource{hs[nxt], ps[nxt], n_dev, 0};
    cur = nxt;
  }
  check_hip(hipStreamSynchronize(s), "sweep sort");
  return cur;
}

void sweep_bounds(const uint64_t *h, const uint64_t *n_dev, uint64_t n_max, uint64_t count,
                  uint32_t *start, hipStream_t s) {
  bucket_bounds_kernel<<<(uint32_t)std::min<uint64_t>(65536, n_max / 256 + 2), 256, 0, s>>>(
      h, n_dev, count, n_max, start);
}

// The per-read sums of hit rows [p0, p0 + n_pos) of a chunk; with a bitmap
// of listed positions the rows of the others are zeroed first, as the sweep
// does.
void sweep_sum_rows(uint32_t *hits, uint32_t row_words, const uint32_t *word_color,
                    const uint64_t *listed, const uint64_t *pos_off, uint64_t n_reads, uint64_t p0,
                    uint64_t n_pos, uint32_t records, uint32_t *counts, uint64_t *partial,
                    uint32_t n_colors, hipStream_t s) {
  uint32_t lp_bits = 0;
  while ((1u << lp_bits) < std::min<uint32_t>(64, row_words)) lp_bits++;
  const uint64_t blocks =
      (uint64_t)((row_words + 63) / 64) * (((n_pos + kSumSpan - 1) / kSumSpan + 3) / 4);
  if (n_pos && listed)
    zero_unlisted_kernel<<<(uint32_t)std::min<uint64_t>(16384, ((n_pos + 63) / 64 + 3) / 4), 256,
                           0, s>>>(hits, row_words, listed, n_pos);
  if (n_pos)
    sweep_sum_kernel<<<(uint32_t)blocks, 256, 0, s>>>(hits, row_words, lp_bits, word_color, pos_off,
                                                      n_reads, p0, n_pos, records, counts, partial,
                                                      n_colors);
}

}  // namespace zg
