// zg_kernels.h -- device-side argument blocks and launchers of the gfx950
// kernels (zg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zg_hash.h"

namespace zg {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Throws with the runtime's message when e is not hipSuccess.
void check_hip(hipError_t e, const char *what);

// One bank resident in HBM, in the plane layout (fingerprint_first).
struct BankDev {
  const uint8_t *data;        // slots * row_bytes (+ padding)
  uint64_t row_bytes;         // device row pitch: the row length rounded up to 16 bytes
  uint64_t plane_bytes;       // ceil(G * b / 8)
  uint32_t groups;            // G
  uint32_t first_merged;      // columns >= this are multi-member groups
  uint32_t hashes;
  uint32_t partition, partitions;
  uint32_t n_tiles;           // row tiles of 64 lanes x field bytes
  SegmentGeometry geo;
  const uint32_t *col_color;  // [G] colour of a singleton column
  const uint32_t *mg_off;     // [G - first_merged + 1] into mg_members
  const uint32_t *mg_members; // colours of merged groups, bit order
  const uint64_t *mg_mask_bit;// [G - first_merged] bit offset of the mask field
  const uint32_t *color_col;  // [n_colors] column of a colour in this bank or kNone
  const uint32_t *color_bit;  // [n_colors] member bit inside its merged group
  uint32_t colmap;            // id shared by banks with the same colour -> column map
};

// Recovery directory: open addressing on the high hash bits, 64-bit entries
// (hash << 32 | posting index), empty = low word kNone. Postings hold
// [count, origin...] with global colour ids; post_col is parallel to post and
// holds each origin's (column map id << 24 | column), kNone where the origin
// is in no bank.
struct RecDev {
  const uint64_t *dir;
  const uint32_t *post;
  const uint32_t *post_col;
  uint32_t cap_log2;          // directory capacity 2^cap_log2
  uint32_t enabled;
  const uint32_t *filter;     // 2^kRecFilterLog2 bits: bits rec_filter_bit(hash, j),
                              // j < filter_probes, set for every directory hash
                              // (a key with any of its bits clear has no postings)
  uint32_t filter_probes;     // 1 .. kRecFilterProbes
};
constexpr uint32_t kRecFilterLog2 = 20;
constexpr uint32_t kRecFilterWords = (1u << kRecFilterLog2) / 32;
constexpr uint32_t kRecFilterProbes = 3;

// Probes of the recovery filter for n directory hashes: the Bloom-filter
// optimum 2^20 ln 2 / n, between 1 and kRecFilterProbes (three probes pass
// about 2 % of absent keys at ~110k hashes, one 10 %; at a million hashes
// more probes pass as many and cost more LDS reads).
inline uint32_t rec_filter_probes(uint64_t n) {
  const double k = n ? (double)(1u << kRecFilterLog2) * 0.6931 / (double)n : kRecFilterProbes;
  return k < 1.5 ? 1 : (k < 2.5 ? 2 : kRecFilterProbes);
}

// Bit j of the recovery filter for a key's abandoned hash: the top 20 bits of
// the hash, then of the hash times two odd constants.
__host__ __device__ inline uint32_t rec_filter_bit(uint32_t ah, uint32_t j) {
  const uint32_t x = j == 0 ? ah : ah * (j == 1 ? 0x9E3779B1u : 0x85EBCA77u);
  return x >> (32 - kRecFilterLog2);
}

enum SeedSource : uint32_t { kSeedKmers = 0, kSeedKeys = 1 };

struct LookupArgs {
  BankDev bank;
  RecDev rec;
  // seeds
  uint32_t source;            // SeedSource
  uint32_t k;
  const uint64_t *codes;      // packed bases, 32 per word (kSeedKmers)
  const uint64_t *invalid;    // invalid-base bits, 64 per word (kSeedKmers)
  const uint64_t *keys;       // explicit keys (kSeedKeys)
  const uint64_t *seg_off;    // [n_reads + 1] read starts in the seed index space
  uint64_t n_reads;
  uint64_t begin, end;        // seed index range of this launch
  uint64_t n_units;           // spans * tile groups
  uint32_t span;              // seed indices per work unit (multiple of 64)
  uint32_t n_tg;              // tile groups per span
  uint32_t *queue;            // work counter(s), zero before launch
  uint32_t slices;            // > 1: units split into that many ranges, one per
                              // block residue (one per XCD), counters at
                              // queue[32 * slice]; at most the grid's block count
  uint64_t seed;
  uint32_t bits;
  uint32_t records;           // 1: per-read counts; 0: per-wave partial sums
  uint32_t *counts;           // [n_reads * n_colors]
  uint64_t *partial;          // [waves * n_colors]
  uint32_t n_colors;
  uint32_t nontemporal;       // 1: row loads bypass cache retention
  uint64_t sampling;          // modimizers: keep k-mers with mixsplit % sampling == 0 (0 = all)
  SpacedRuns spaced;          // spaced seed applied to each k-mer (kSeedKmers; n = 0: none)
  // findere: instead of counting, each seed's hits go to its row of fd_hits
  // (fd_row_words 32-bit words per seed index). This bank's region starts at
  // word fd_off: bit = device column for singleton columns, then bit
  // fd_member_bit + member slot (index into mg_members) for merged groups.
  uint32_t *fd_hits;          // nullptr: count as usual
  uint64_t fd_row_words;
  uint32_t fd_off;
  uint32_t fd_member_bit;
};

struct KernelConfig {
  uint32_t bits;
  uint32_t tiles_per_wave;  // 1 or 2
  uint32_t prefetch;        // 1, 2 or 4
  uint32_t blocks;          // grid size (256-thread blocks)
};

// Field bytes each lane reads per row and tile for fingerprint width b.
inline uint32_t lane_bytes_for_bits(uint32_t b) {
  return b == 1 ? 2 : b == 2 ? 4 : b == 4 ? 8 : 16;
}
inline uint32_t fields_per_lane(uint32_t b) { return lane_bytes_for_bits(b) * 8 / b; }

// Launchers (all asynchronous on `stream`).
void launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);
void launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s);
// ASCII reads -> packed codes + invalid bits in position space (one invalid
// separator after each read). pos_off[r] = read_off[r] + r.
void launch_pack(const uint8_t *seq, const uint64_t *read_off, const uint64_t *pos_off,
                 uint64_t n_reads, uint64_t positions, uint64_t *codes, uint64_t *invalid,
                 hipStream_t s);
// Valid k-mer windows per read (atomically added into totals).
// sampling > 0 counts only modimizers (mixsplit(canonical, seed) % sampling == 0).
void launch_count_kmers(const uint64_t *invalid, const uint64_t *codes, uint64_t seed,
                        uint64_t sampling, const uint64_t *pos_off, uint64_t n_reads,
                        uint64_t positions, uint32_t k, uint64_t *totals, hipStream_t s);
// Open-closed mod-minimizer selection: per window start, flag + offset, and
// per-block and per-read counts.
void launch_ocm_select(const uint64_t *codes, const uint64_t *invalid, const uint64_t *pos_off,
                       uint64_t n_reads, uint64_t positions, uint32_t k, uint32_t m,
                       uint64_t seed, const uint64_t *inner_table, uint8_t *flag,
                       uint8_t *delta, uint32_t *block_counts, uint64_t *totals,
                       hipStream_t s);
void launch_ocm_compact(const uint64_t *codes, const uint8_t *flag, const uint8_t *delta,
                        const uint64_t *block_off, uint64_t positions, uint32_t m,
                        const SpacedRuns &spaced, uint64_t *keys, uint64_t *key_pos,
                        hipStream_t s);
// Findere reconstruction over a region of hit rows (findere::Scorer,
// src/findere.rs:74-104): colour bit_color[j] of region bit j scores a
// position when that bit is set in the z + 1 rows ending there. Adds into
// counts[r * n_colors + colour] (records) or partial[colour].
// Hit rows are indexed by position - p_base (z = 0 counts every hit).
void launch_findere_scan(const uint32_t *hits, uint64_t row_words, uint32_t region_off,
                         uint32_t region_bits, const uint32_t *bit_color,
                         const uint64_t *pos_off, uint64_t n_reads, uint64_t positions,
                         uint32_t z, uint32_t records, uint32_t *counts, uint64_t *partial,
                         uint32_t n_colors, hipStream_t s, uint64_t p_base = 0);
// Padding, in words, after the invalid-base bits of `positions` positions so
// that a window of `window` bases starting at any position stays in bounds.
inline uint64_t invalid_words(uint64_t positions, uint64_t window) {
  return (positions + 63) / 64 + 4 + (window + 63) / 64;
}
inline uint64_t code_words(uint64_t positions, uint64_t window) {
  return (positions + 31) / 32 + 4 + (window + 31) / 32;
}
// Long k-mers (k > 32): rolling polynomial hashes of both orientations
// (scan_long_kmers, src/lib.rs:683-730).
struct LongKmerParams {
  uint64_t power;    // base^(k-1)
  uint64_t inverse;  // base^-1 mod 2^64
  uint64_t domain;   // splitmix64(k ^ 0x6b6d65722d763031)
  uint32_t k;
};
LongKmerParams long_kmer_params(uint32_t k);
// Per 64 positions: keep bits of the valid windows (modimizers: sampled
// ones), their count, and the key of each kept window at pos_key[p].
void launch_long_kmer_keys(const uint64_t *codes, const uint64_t *invalid, uint64_t positions,
                           const LongKmerParams &lp, uint64_t seed, uint64_t sampling,
                           uint64_t *keep, uint32_t *counts, uint64_t *pos_key, hipStream_t s);
// keys[off[w] + rank] = pos_key[p], key_win[...] = p for every kept bit p.
void launch_compact_kept(const uint64_t *keep, const uint64_t *off, const uint64_t *pos_key,
                         uint64_t positions, uint64_t *keys, uint64_t *key_win, hipStream_t s);
// Randstrobemers: candidate bitmaps (64 positions per word) with per-word
// counts, their compaction into (start, strobe hash, run id) in position
// order, and the linker for both query orientations (keys[i] forward,
// keys[n + i] reverse complement).
void launch_popc_words(const uint64_t *words, uint64_t n, uint32_t *counts, hipStream_t s);
void launch_strobe_kmer_candidates(const uint64_t *invalid, uint64_t positions, uint32_t k,
                                   uint64_t *cand, uint32_t *counts, hipStream_t s);
void launch_strobe_mark_selected(const uint8_t *flag, const uint8_t *delta, uint64_t positions,
                                 uint64_t *cand, hipStream_t s);
void launch_strobe_compact(const uint64_t *cand, const uint64_t *cand_off, const uint64_t *invalid,
                           const uint64_t *inv_off, const uint64_t *codes, uint64_t positions,
                           uint32_t len, const SpacedRuns &spaced, uint64_t *cpos,
                           uint64_t *chash, uint64_t *crun, hipStream_t s);
void launch_strobe_link(const uint64_t *cpos, const uint64_t *chash, const uint64_t *crun,
                        uint64_t n, uint64_t lo, uint64_t hi, uint64_t *keys, hipStream_t s);
void launch_strobe_offsets(uint64_t *key_off, uint64_t n_reads, uint64_t total, hipStream_t s);
// counts[i] = max(counts[i], counts[n + i]) for i < n (the two orientations).
void launch_strand_max(uint32_t *counts, uint64_t n, hipStream_t s);
// Homopolymer compression of packed reads: keep flags per 64 positions, then
// (after an exclusive scan of counts into keep_off) the placement of kept
// bases into zeroed codes / all-ones invalid.
void launch_hpc_keep(const uint8_t *seq, const uint64_t *read_off, const uint64_t *pos_off,
                     uint64_t n_reads, uint64_t positions, uint64_t *keep_bits,
                     uint32_t *counts, hipStream_t s);
void launch_hpc_place(const uint8_t *seq, const uint64_t *read_off, const uint64_t *pos_off,
                      uint64_t n_reads, uint64_t positions, const uint64_t *keep_bits,
                      const uint64_t *keep_off, uint64_t *codes, uint64_t *invalid,
                      hipStream_t s);
uint32_t ocm_block_positions();
// Exclusive prefix sums (zg_scan.hip): block sums of 4096-element tiles,
// scanned recursively, then each tile scanned from its block offset. The
// caller provides the temporary buffer (the block-sum levels); in may equal
// out for the u64 form.
size_t scan_temp_bytes_u64(uint64_t n);
void launch_exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, void *temp,
                               size_t temp_bytes, hipStream_t s);
void launch_exclusive_scan_u32_to_u64(const uint32_t *in, uint64_t *out, uint64_t n,
                                      void *temp, size_t temp_bytes, hipStream_t s);
size_t scan_temp_bytes_u32_to_u64(uint64_t n);
// The lookup.
void launch_lookup(const LookupArgs &a, const KernelConfig &c, hipStream_t s);
// Sum per-wave partials into agg[n_colors] (added).
void launch_reduce_partials(const uint64_t *partial, uint64_t waves, uint32_t n_colors,
                            uint64_t *agg, hipStream_t s);
// Occupancy-derived number of resident 256-thread blocks per CU for a config.
int lookup_blocks_per_cu(const KernelConfig &c);
// n random whole-row reads of row_bytes (512..8192, powers of two) over
// data[0, bytes); false for an unsupported row size.
bool launch_gather_probe(const uint8_t *data, uint64_t bytes, uint32_t row_bytes, uint64_t n,
                         uint32_t blocks, uint32_t *sink, hipStream_t s);
// The H rows of n_keys random keys under a bank's PositionHash geometry
// (whole rows of row_bytes 512, 1024 or 2048); false for other row sizes.
bool launch_gather_keys(const uint8_t *data, uint64_t row_bytes, SegmentGeometry geo, uint32_t H,
                        uint64_t n_keys, uint32_t blocks, uint32_t *sink, hipStream_t s);

}  // namespace zg

namespace zg {
// key_off[r] = first key whose window start is >= pos_off[r] (r <= n_reads).
void launch_segment_bounds(const uint64_t *key_win, uint64_t n_keys, const uint64_t *pos_off,
                           uint64_t n_reads, uint64_t *key_off, hipStream_t s);
// totals[r] = off[r+1] - off[r]
void launch_diff(const uint64_t *off, uint64_t n, uint64_t *totals, hipStream_t s);
// agg[c] += sum over reads of counts[r * n_colors + c]
void launch_colsum(const uint32_t *counts, uint64_t n_reads, uint32_t n_colors, uint64_t *agg,
                   hipStream_t s);
}  // namespace zg
