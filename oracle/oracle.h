/*
 * oracle.h -- CPU restatement of the zorindex query path (test infrastructure).
 *
 * This library is the parity oracle and the timed CPU baseline. It is NOT part
 * of the product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only as the checker. The product path
 * (libzindex_amd.so) never links or calls it.
 *
 * Every function restates one piece of the reference (Rust crate zorindex);
 * the citation next to each declaration names the file:line it follows.
 *
 * Pinning status (what checks each function against something other than
 * itself) is listed per function in oracle.c.
 */
#ifndef ZINDEX_ORACLE_H
#define ZINDEX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0 hash primitives --------------------------------------------- */
/* utils/ocm-minimizers/src/lib.rs:40-46 */
uint64_t zo_splitmix64(uint64_t x);
/* utils/ocm-minimizers/src/lib.rs:35-38 */
uint64_t zo_mixsplit(uint64_t key, uint64_t seed);
/* utils/ocm-minimizers/src/lib.rs:7-15; returns 0..3, or -1 for a non-ACGT byte */
int zo_encode_base(uint8_t base);
/* utils/ocm-minimizers/src/lib.rs:17-23; returns 1 and stores the value, or 0 */
int zo_encode_full_kmer(const uint8_t *seq, size_t len, uint64_t *out);
/* utils/ocm-minimizers/src/lib.rs:25-34 */
uint64_t zo_canonical_encoded(uint64_t encoded, int k);
/* src/power_merge/storage.rs:3214-3216 */
uint64_t zo_fingerprint(uint64_t key, uint64_t seed, int bits);
/* src/power_merge/abandoned.rs:8-10 */
uint32_t zo_abandoned_hash(uint64_t key, uint64_t seed);
/* src/power_merge/storage.rs:99-101 */
uint64_t zo_partition_id(uint64_t key, uint64_t seed, uint64_t partitions);
/* src/power_merge/storage.rs:3452-3481; segment_length 0 = legacy (None) */
void zo_positions(uint64_t key, uint64_t seed, uint64_t cells, int hashes,
                  uint64_t segment_length, uint64_t *out);
/* src/power_merge/storage.rs:3525-3538 */
uint64_t zo_get_bits(const uint8_t *row, size_t offset, size_t width);
/* src/power_merge/storage.rs:3495-3522 (row must be zero where written) */
void zo_put_bits(uint8_t *row, size_t row_len, size_t offset, size_t width,
                 uint64_t value);

/* ---- L1 seed extraction ----------------------------------------------- */
/* src/lib.rs:645-681 (k<=32) and 687-731 (k>32). Returns the number of keys;
 * out may be NULL to count only. out must hold len keys. */
size_t zo_scan_kmers(const uint8_t *seq, size_t len, int k, uint64_t *out);
/* utils/ocm-minimizers/src/implementation.rs:108-236 (portable rolling path).
 * pos_out and key_out hold up to len entries; either may be NULL. */
size_t zo_scan_ocm(const uint8_t *seq, size_t len, int k, int m, uint64_t seed,
                   uint64_t *pos_out, uint64_t *key_out);
/* src/open_closed_minimizers.rs:43-85: the reference's brute-force test
 * oracle, evaluated on one valid (all-ACGT) run. */
size_t zo_scan_ocm_bruteforce(const uint8_t *run, size_t len, int k, int m,
                              uint64_t seed, uint64_t *pos_out,
                              uint64_t *key_out);

/* ---- L3 lookup ---------------------------------------------------------- */
typedef struct zo_bank {
  const uint8_t *data;        /* slots * ceil(row_bits/8) bytes */
  uint64_t slots;
  uint64_t row_bits;
  uint32_t n_groups;
  const uint32_t *member_off; /* n_groups + 1 offsets into members */
  const uint32_t *members;    /* color ids, in group order */
  const uint32_t *width;      /* per group */
  const uint64_t *offset;     /* per group */
  int fingerprint_first;      /* 1 = plane layout (ZORPOW3), 0 = legacy */
  uint32_t partition;         /* routing; partitions == 1 without routing */
  uint32_t partitions;
} zo_bank;

typedef struct zo_recovery {
  uint64_t n;                 /* distinct 32-bit hashes, ascending */
  const uint32_t *hashes;
  const uint64_t *off;        /* n + 1 offsets into origins */
  const uint32_t *origins;    /* ascending per hash */
} zo_recovery;

typedef struct zo_index {
  int k;
  int hashes;
  int bits;
  int mode;                   /* 0 = kmers, 1 = minimizers (open-closed mod), 2 = modimizers */
  int m;                      /* minimizer size */
  uint64_t seed;
  uint64_t segment_length;    /* 0 = legacy None */
  uint32_t n_colors;
  const zo_bank *banks;
  uint32_t n_banks;
  const zo_recovery *recovery; /* NULL = no recovery */
  uint64_t sampling;          /* modimizer sampling (mode 2) */
  int hpc;                    /* homopolymer compression before seeding */
  uint32_t seed_width;        /* spaced seed width (0 = none) */
  uint64_t seed_mask;         /* 3 per retained base, first base most significant */
  uint32_t findere_z;         /* mode 3: z (s-mers of k - z bases are indexed) */
  uint64_t strobe_min;        /* randstrobemers: inclusive distance bounds */
  uint64_t strobe_max;        /*   (strobe_max 0 = no randstrobemers) */
} zo_index;

/* strobe_hash (src/anchors.rs:134-151): XXH64 of one u64, seed 0. */
uint64_t zo_strobe_hash(uint64_t value);

/* Linker (src/anchors.rs:185-256) over n candidates of one run in position
 * order: keys[i] = hash[i]/2 + hash[partner]/2, the partner being the later
 * candidate at distance in [min_d, max_d] with the fewest differing low eight
 * hash bits (first on ties), or i itself. */
void zo_link_strobes(const uint64_t *pos, const uint64_t *hash, size_t n, uint64_t min_d,
                     uint64_t max_d, uint64_t *keys);

/* findere::Scorer::sequence (src/findere.rs:74-104): per run of valid bases
 * of at least k, every canonical s-mer (s = k - z) is looked up through
 * `lookup` (which fills hits[n_colors], zeroed before each call); a colour
 * scores a k-mer when its z + 1 consecutive s-mers all hit. Adds to scores,
 * *total (k-mer windows) and *smer_queries. */
typedef void (*zo_hits_fn)(void *ctx, uint64_t key, uint64_t *hits);
void zo_findere_sequence(const uint8_t *seq, size_t len, int k, int z, uint32_t n_colors,
                         zo_hits_fn lookup, void *ctx, uint64_t *scores, uint64_t *total,
                         uint64_t *smer_queries);

/* homopolymer_compress (src/lib.rs:209-222); returns the output length. */
size_t zo_homopolymer_compress(const uint8_t *seq, size_t len, uint8_t *out);

/* SpacedSeed::project_portable and ::canonical (src/anchors.rs:67-90). */
uint64_t zo_spaced_project(uint64_t encoded, uint32_t width, uint64_t mask);
uint64_t zo_spaced_canonical(uint64_t encoded, uint32_t width, uint64_t mask);

/* src/power_merge/storage.rs:2761-2925 (lookup_selected, ranges = None),
 * with WeightedScores (3123-3140). Adds weight to scores[color]. */
void zo_lookup(const zo_index *idx, uint64_t key, uint64_t weight,
               uint64_t *scores);

/* Feature extraction for one read: src/lib.rs:574-642 (kmers, OCM). */
size_t zo_read_features(const zo_index *idx, const uint8_t *seq, size_t len,
                        uint64_t *out);

/* Per-read counts: src/power_merge/query_stream.rs:163-452 with
 * record_query.rs:337-404 (sort, fold duplicates into weights, look up).
 * counts: n_reads * n_colors u64 (row-major); totals: n_reads. Reads are
 * seq[read_off[i] .. read_off[i+1]). threads <= 0 means 1. */
int zo_query_reads(const zo_index *idx, const uint8_t *seq,
                   const uint64_t *read_off, uint64_t n_reads,
                   uint64_t *counts, uint64_t *totals, int threads);

/* Aggregate counts over all reads (query_stream.rs:281-299):
 * agg[n_colors], returns total features. */
uint64_t zo_query_aggregate(const zo_index *idx, const uint8_t *seq,
                            const uint64_t *read_off, uint64_t n_reads,
                            uint64_t *agg, int threads);

/* Same lookups over an explicit key list (keys already extracted), used by
 * the CPU baseline to time the lookup phase alone. */
int zo_lookup_keys(const zo_index *idx, const uint64_t *keys,
                   const uint64_t *key_off, uint64_t n_reads,
                   uint64_t *counts, int threads);

/* The same counts by the reference's optimised lookup (record_query.rs:
 * 337-404: sorted keys folded into weights, blocks of 16 distinct keys, rows
 * prefetched 4 keys ahead; storage.rs:3309-3350: AVX2 screen of byte-wide
 * fingerprint planes), reads spread over the threads: the timed CPU
 * baseline. */
int zo_lookup_keys_fast(const zo_index *idx, const uint64_t *keys,
                        const uint64_t *key_off, uint64_t n_reads,
                        uint64_t *counts, int threads);

/* Fingerprint screen (storage.rs:3241-3350, packed_fingerprints.rs:7-126):
 * writes hit group indices of XOR(rows) == fp, returns their number. */
size_t zo_screen_rows(const uint8_t *const *rows, int n_rows, int bits,
                      uint64_t fp, size_t count, uint32_t *hits);
/* The materialized-row screen used as its test oracle (storage.rs:3384-3450). */
size_t zo_screen_materialized(const uint8_t *row, int bits, uint64_t fp,
                              size_t count, uint32_t *hits);

#ifdef __cplusplus
}
#endif
#endif
