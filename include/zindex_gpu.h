/*
 * zindex_gpu.h -- C ABI of libzindex_amd.so, the MI355X (gfx950) query path
 * for zorindex indexes.
 *
 * The reference (Rust crate `zorindex`) has no FFI; its drop-in boundary is the
 * library API re-exported at src/power_merge/mod.rs:45-49 plus the CLI
 * (src/bin/zorindex.rs). Each entry point below names the reference function
 * it replaces. Signatures carry plain pointers and sizes only; every buffer is
 * caller-owned unless stated otherwise, and only opaque handles cross the ABI.
 *
 * Errors: every int-returning call returns ZG_OK (0) or a negative ZG_ERR_*
 * code and leaves a message in the calling thread's zg_last_error().
 * No C++ exception crosses the boundary.
 *
 * Threading: one handle per host thread, or external serialization. Device
 * state owned by an index is read-only during queries.
 */
#ifndef ZINDEX_GPU_H
#define ZINDEX_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZG_OK 0
#define ZG_ERR_INVALID (-1)     /* bad argument or failed validation */
#define ZG_ERR_IO (-2)          /* file system or decoder failure */
#define ZG_ERR_FORMAT (-3)      /* index/manifest/bank format violation */
#define ZG_ERR_DEVICE (-4)      /* HIP runtime failure or no usable GPU */
#define ZG_ERR_NOMEM (-5)       /* host or device allocation failed */
#define ZG_ERR_UNSUPPORTED (-6) /* valid index, mode this library lacks */
#define ZG_ERR_BUDGET (-7)      /* HBM admission budget exceeded */

/* Query flags. */
#define ZG_QUERY_RECORDS 0x1u     /* per-read counts (else aggregate only) */
#define ZG_QUERY_NO_RECOVERY 0x2u /* ignore the abandoned-key recovery map */

/* Query memory modes (QueryMemory, src/power_merge/mod.rs:52-59). */
#define ZG_MEMORY_ALL 0u      /* every subindex resident in HBM */
#define ZG_MEMORY_SUBINDEX 1u /* subindexes resident while they fit the HBM
                                 budget; the rest stream through HBM one at a
                                 time during each query */

/* Thread-local message for the last failing call on this thread. */
const char *zg_last_error(void);
/* Library version string. */
const char *zg_version(void);

/* ------------------------------------------------------------------------ */
/* Index handle                                                             */
/* ------------------------------------------------------------------------ */
typedef struct zg_index zg_index;

typedef struct zg_index_meta {
  uint32_t version;
  uint32_t k;
  uint32_t hashes;
  uint32_t fingerprint_bits;
  uint32_t index_mode;        /* 0 kmers, 1 minimizers, 2 modimizers, 3 findere */
  uint32_t minimizer_size;
  uint32_t minimizer_scheme;  /* 0 simd, 1 open-closed-mod */
  uint32_t n_colors;
  uint32_t n_banks;
  uint32_t n_classes;         /* distinct (component, source_slots) groups */
  uint64_t seed;
  uint64_t segment_length;    /* 0 when the manifest has none (legacy) */
  int32_t abandoned_recovery;
  uint32_t reserved;
  uint64_t abandoned_hashes;
  uint64_t abandoned_origins;
  uint64_t payload_bytes;     /* sum of bank payloads */
} zg_index_meta;

/* Replaces Index::load_mode(dir, recover, None, metadata_only=true)
 * (src/power_merge/storage.rs:2273-2637): parses manifest.json and applies the
 * full validation set before any payload is touched. */
int zg_index_open(const char *dir, zg_index **out);
/* Releases host and device memory of the handle. */
int zg_index_close(zg_index *idx);
int zg_index_get_meta(const zg_index *idx, zg_index_meta *out);
/* Colour name and unique_kmers (storage.rs:57-63); *name stays valid for the
 * lifetime of idx. */
int zg_index_color(const zg_index *idx, uint32_t color, const char **name,
                   uint64_t *unique_kmers);

/* Replaces load_modern_bank + bank_io::read_parallel (storage.rs:2232-2256,
 * bank_io.rs:122-291) and QueryRecovery/load_abandoned (query_stream.rs:8-88,
 * storage.rs:2136-2181): decodes every bank and (unless no_recovery) the
 * recovery map, then makes them resident in HBM. Fails with ZG_ERR_BUDGET
 * when the payload exceeds hbm_budget_bytes (0 = no budget). */
int zg_index_load(zg_index *idx, int recover, uint64_t hbm_budget_bytes);
/* zg_index_load with a query memory mode. Bank files are decoded frame by
 * frame in parallel into page-locked chunks that are copied to HBM while the
 * next chunk decodes (bank_io.rs:235-291); host copies are not kept. With
 * ZG_MEMORY_SUBINDEX and an index larger than the budget, the banks that do
 * not fit are decoded and streamed in per query (the reference's default
 * subindex-at-a-time mode); the budget must hold the largest bank beside the
 * resident ones. */
int zg_index_load_ex(zg_index *idx, int recover, uint64_t hbm_budget_bytes,
                     uint32_t memory);
/* Host copy of one decoded bank payload (valid after zg_index_load, or after
 * an in-memory build). */
int zg_index_bank_payload(const zg_index *idx, uint32_t bank,
                          const uint8_t **data, uint64_t *bytes);
/* manifest.json text of the index (storage.rs:133-258 schema); valid for the
 * lifetime of idx. */
int zg_index_manifest(const zg_index *idx, const char **json, uint64_t *len);
/* Host recovery map in memory (after zg_index_load with recovery or an
 * in-memory build): n hashes ascending, off[n+1] into origins
 * (QueryRecovery entries, query_stream.rs:8-46). */
int zg_index_recovery(const zg_index *idx, const uint32_t **hashes,
                      const uint64_t **off, const uint32_t **origins,
                      uint64_t *n);

/* ------------------------------------------------------------------------ */
/* Queries on host reads                                                     */
/* ------------------------------------------------------------------------ */
/* Replaces query_ordinary (src/power_merge/query_stream.rs:163-452) with
 * record_query::Worker::query_prepared (record_query.rs:302-404):
 * reads are seq[read_off[i] .. read_off[i+1]) (ASCII, any bytes; non-ACGT
 * splits runs). Outputs (each may be NULL):
 *   totals[n_reads]            seeds per read (queried_seeds / valid_kmers)
 *   counts[n_reads * n_colors] per-read per-colour matches (ZG_QUERY_RECORDS)
 *   aggregate[n_colors]        matches summed over all reads
 * Matches are count[c] += [filter hit(c) OR c recovered] per seed occurrence,
 * the exact semantics of Index::lookup_selected (storage.rs:2761-2925). */
int zg_query_reads(zg_index *idx, const uint8_t *seq, const uint64_t *read_off,
                   uint64_t n_reads, uint32_t flags, uint64_t *totals,
                   uint32_t *counts, uint64_t *aggregate);

/* No counterpart in the reference: zg_query_reads for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). It copies host buffers and synchronizes, so it is not for
 * stream capture. Same
 * arguments, results and errors as zg_query_reads (bit-identical outputs);
 * every device buffer comes from one static 64 MiB array per GPU, allocated
 * when the library's code is loaded and held until the process ends. A call
 * whose working set exceeds it (about 1.4 bytes per base plus 24 bytes per
 * read, plus 4 bytes per read and colour with ZG_QUERY_RECORDS; minimizers add
 * about 2 bytes per base and 16 per key) fails with ZG_ERR_INVALID before any
 * GPU work. Covers k-mers (k <= 32), minimizers and modimizers without
 * homopolymer compression, every subindex resident; other indexes get
 * ZG_ERR_UNSUPPORTED. Calls that overlap on one GPU share the array and race:
 * serialize them (one host thread per GPU). */
int zg_query_reads_noalloc(zg_index *idx, const uint8_t *seq,
                           const uint64_t *read_off, uint64_t n_reads,
                           uint32_t flags, uint64_t *totals, uint32_t *counts,
                           uint64_t *aggregate);

/* Seeds the GPU selects for each read (scan_index_features, src/lib.rs:574;
 * open_closed_minimizers::scan_positions, utils/ocm-minimizers/src/lib.rs:4):
 * key_off[n_reads+1]; keys and positions (offset within the read) hold up to
 * cap entries. Returns ZG_ERR_INVALID if cap is too small (key_off[n_reads]
 * then holds the required size). */
int zg_extract_seeds(zg_index *idx, const uint8_t *seq,
                     const uint64_t *read_off, uint64_t n_reads,
                     uint64_t *key_off, uint64_t *keys, uint64_t *positions,
                     uint64_t cap);

/* ------------------------------------------------------------------------ */
/* CLI-equivalent commands (byte-identical stdout to the reference)          */
/* ------------------------------------------------------------------------ */
/* query_command_with_memory (storage.rs:3731-3767) via
 * query_stream::query_aggregate (query_stream.rs:864-946). */
int zg_query_command(const char *dir, const char *query, double min_ratio,
                     int recover, int out_fd, int err_fd);
/* query_records_command_with_memory (storage.rs:3626-3720); format 0 = tsv,
 * 1 = compact (record_output.rs:5-63). */
int zg_query_records_command(const char *dir, const char *query,
                             double min_ratio, int recover,
                             uint64_t batch_size, int format, int out_fd,
                             int err_fd);
/* The commands above with a query memory mode and an HBM budget (0 = none). */
int zg_query_command_ex(const char *dir, const char *query, double min_ratio,
                        int recover, uint32_t memory, uint64_t hbm_budget,
                        int out_fd, int err_fd);
int zg_query_records_command_ex(const char *dir, const char *query,
                                double min_ratio, int recover,
                                uint64_t batch_size, int format,
                                uint32_t memory, uint64_t hbm_budget,
                                int out_fd, int err_fd);
/* info_command (storage.rs:3901-4092). */
int zg_info_command(const char *dir, int out_fd);

/* ------------------------------------------------------------------------ */
/* Index construction in the reference on-disk format                        */
/* ------------------------------------------------------------------------ */
typedef struct zg_build_config {
  uint32_t k;                  /* default 31 */
  uint32_t hashes;             /* 2..32, default 4 */
  uint32_t fingerprint_bits;   /* 1,2,4,8,16,32, default 8 */
  uint32_t index_mode;         /* 0 kmers, 1 minimizers (open-closed mod), 2 modimizers */
  uint32_t minimizer_size;     /* m for minimizers */
  uint32_t segment_length;     /* default 64 */
  uint64_t seed;               /* default 69 */
  double slot_scale;           /* default 1.0 */
  uint64_t max_subindex_size;  /* power of two, default 8 GiB */
  uint32_t merge_group_size;   /* 1 = singleton groups; n>1 merges n
                                  consecutive datasets per group */
  uint32_t threads;            /* 0 = all host cores */
  int32_t drop_abandoned;      /* 1 = no recovery map */
  uint32_t modimizer_sampling; /* index_mode 2 (modimizers): keep canonical
                                  k-mers with mixsplit % sampling == 0;
                                  default 16 */
  uint32_t device;             /* 1 = peel on the GPU: parallel peeling rounds,
                                  back-substitution and verification as HIP
                                  kernels; a 2-core left by peeling is broken
                                  on the host by the same rule as the CPU build
                                  (the core is unique, so the abandoned keys
                                  are the same); 0 = CPU */
  uint32_t homopolymer_compression; /* 1 = collapse runs of one base (case-
                                       insensitive; other bytes kept as
                                       breaks) before seeding, at build and
                                       query (src/lib.rs:209-222) */
  uint32_t findere_z;          /* index_mode 3 (findere): index (k - z)-mers and
                                  count a k-mer when its z + 1 s-mers all hit;
                                  default 10 */
  char seed_pattern[32];       /* spaced seed: a symmetric '#'/'_' pattern of
                                  width k (m for minimizers), NUL-terminated;
                                  "" = none (src/anchors.rs:10-90) */
  uint32_t randstrobemers;     /* 1 = link two k-mers or selected minimizers
                                  into randstrobemers (src/anchors.rs:136-256) */
  uint32_t strobe_min;         /* inclusive strobe-start distance bounds;
                                  0 = the strobe length (k or m) and three */
  uint32_t strobe_max;         /*   times it */
} zg_build_config;

/* Fills cfg with the reference CLI defaults (src/bin/zorindex.rs:55-159). */
void zg_build_config_default(zg_build_config *cfg);

/* power_merge::build (storage.rs:500-882) with the naive planner and
 * ColumnOrder::Width: reads a file-of-files ("name<TAB>path" or "path" lines,
 * src/lib.rs read_fof) and writes manifest.json, bank files and
 * abandoned.zormap.zst into out_dir (which must not exist). */
int zg_build_command(const zg_build_config *cfg, const char *fof,
                     const char *out_dir, int err_fd);

/* fuse_many_indexes (src/power_merge/fusion.rs:233-480): concatenates two
 * or more compatible indexes into out_dir (which must not exist) without
 * rereading their datasets: banks copied as components 1..n, dataset ids of
 * later inputs offset, recovery maps merged, manifest version 10. */
int zg_fusion_command(const char *const *inputs, uint32_t n_inputs,
                      const char *out_dir, int err_fd);

typedef struct zg_builder zg_builder;
int zg_builder_create(const zg_build_config *cfg, zg_builder **out);
/* Adds one dataset: records are seq[rec_off[i] .. rec_off[i+1]). The dataset
 * is extracted, deduplicated and peeled immediately; only its cells are kept. */
int zg_builder_add_color(zg_builder *b, const char *name, const uint8_t *seq,
                         const uint64_t *rec_off, uint64_t n_records);
/* Packs banks and returns an index usable without touching disk
 * (zg_index_load still uploads it). The builder is consumed. */
int zg_builder_finish(zg_builder *b, zg_index **out);
/* Writes the built index to out_dir in the reference format. */
int zg_index_write(const zg_index *idx, const char *out_dir);
int zg_builder_destroy(zg_builder *b);

/* ------------------------------------------------------------------------ */
/* Device-resident batches (persistent kernel, HIP graph, plain path)         */
/* ------------------------------------------------------------------------ */
typedef struct zg_batch zg_batch;

typedef struct zg_tuning {
  uint32_t span;           /* seed positions per work unit (multiple of 64) */
  uint32_t tiles_per_wave; /* row tiles (64 lanes x field bytes) per wave */
  uint32_t prefetch;       /* seeds whose rows are in flight per wave: 1, 2,
                              4; also 8, 16 for fingerprints of <= 4 bits */
  uint32_t waves_per_cu;   /* persistent waves per compute unit (0 = max) */
  uint32_t nontemporal;    /* 1: row gathers use non-temporal loads */
  uint32_t sweep;          /* 1: seeds are looked up in PositionHash segment
                              order (a radix sort per chunk of reads and size
                              class) from segment windows staged in LDS, each
                              seed's hits into a hit row at its position, the
                              rows summed per read; same results as 0.
                              Ignored (read order) for findere indexes,
                              streamed subindexes and rows too wide for LDS */
  uint32_t graph;          /* with sweep: a batch's sweep (every chunk's radix
                              passes, staged lookups and sums) is captured
                              as one HIP graph after its first run and
                              replayed as one launch; its phase times are
                              then not measured (sweep_*_ms = -1) */
} zg_tuning;

typedef struct zg_batch_stats {
  uint64_t n_reads;
  uint64_t bases;
  uint64_t seeds;            /* seed occurrences over all reads */
  uint64_t lookups;          /* seed x bank probes executed */
  uint64_t algorithmic_bytes;/* bytes the lookups must read, per full run */
  double last_lookup_ms;     /* HIP-event time of the last zg_batch_run lookups */
  double last_seed_ms;       /* HIP-event time of packing + seed extraction */
  double last_total_ms;      /* HIP-event time of the whole last run */
  zg_tuning tuning;
  uint64_t sweep_folded;     /* last sorted-sweep run: seed occurrences answered
                                from the occurrence before them in bank order
                                (equal keys, the per-read fold); 0 otherwise */
  uint64_t sweep_launches;   /* its chunks of whole reads */
  double sweep_lookup_ms;    /* HIP-event time of its staged lookups */
  double sweep_sort_ms;      /* ... of its radix passes and segment bounds */
  double sweep_sum_ms;       /* ... of its per-read sums */
  uint64_t sweep_stage_launches;    /* staged lookup launches (chunks x banks) */
  uint64_t sweep_algorithmic_bytes; /* bytes its staged lookups must move per
                                       run: every bank row read once per
                                       chunk, each seed's 8-byte first hash and
                                       4-byte position read once and its hit
                                       row written once */
  uint64_t sweep_hit_row_bytes;     /* hit-row bytes per position */
  uint64_t sweep_plane_banks;       /* banks its lookups read as bit planes (a
                                       copy of the bank made on first use: per
                                       32 fields, b words of one bit each) */
  double sweep_recovery_ms;         /* HIP-event time of its recovery passes
                                       (bit-plane classes; their postings
                                       ORed into the hit rows after the
                                       lookups); not in sweep_lookup_ms */
} zg_batch_stats;

/* Uploads reads (ASCII) to HBM, packs them to 2 bits plus an invalid-base
 * mask and extracts seeds on the device. */
int zg_batch_create(zg_index *idx, const uint8_t *seq,
                    const uint64_t *read_off, uint64_t n_reads, uint32_t flags,
                    zg_batch **out);
/* Runs the batched persistent lookup over every resident bank (replayed from
 * a captured HIP graph after the first call). */
int zg_batch_run(zg_index *idx, zg_batch *batch);
/* The plain baseline: one lookup launch per read and bank, reads
 * [first, first + count). */
int zg_batch_run_plain(zg_index *idx, zg_batch *batch, uint64_t first,
                       uint64_t count);
/* Re-extracts seeds (pack + seed kernels) on the device; timed separately. */
int zg_batch_reseed(zg_index *idx, zg_batch *batch);
int zg_batch_results(zg_batch *batch, uint64_t *totals, uint32_t *counts,
                     uint64_t *aggregate);
int zg_batch_get_stats(const zg_batch *batch, zg_batch_stats *out);
int zg_batch_destroy(zg_batch *batch);

/* Measurement, no reference counterpart: after a sorted-sweep run, writes one
 * whole hit row at the position of every seed of the batch's last chunk, in
 * the sweep's order (segment, then position), each row by neighbouring lanes
 * of one store instruction, with no lookups, and reports the rate in GB/s:
 * the ceiling of the lookups' hit-row stores. Overwrites the batch's hit rows
 * (not its results). ZG_ERR_INVALID if no sweep ran. */
int zg_batch_store_probe(zg_batch *batch, double *gbs);

int zg_index_set_tuning(zg_index *idx, const zg_tuning *tuning);
int zg_index_get_tuning(const zg_index *idx, zg_tuning *out);
/* Times candidate tunings on the batch and keeps the fastest. */
int zg_autotune(zg_index *idx, zg_batch *batch, zg_tuning *chosen);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of query requests over a resident index      */
/* ------------------------------------------------------------------------ */
/* The reference has no server; its query driver (query_ordinary,
 * query_stream.rs:163-452) answers one file per process. A server keeps the
 * index resident and answers requests as they arrive. Each of `slots` worker
 * threads owns a HIP stream and device buffers for one batch; whenever a slot
 * is free it admits the queued requests, in arrival order, that fit its
 * capacity, so batches formed while others are on the GPU overlap with them
 * (copies, seeding and the tail of one persistent lookup with the next). A
 * request's counts do not depend on what it was batched with. */
typedef struct zg_server zg_server;

typedef struct zg_server_config {
  uint32_t slots;         /* batches in flight (one stream each); default 2 */
  uint32_t max_requests;  /* requests admitted into one batch; default 256 */
  uint64_t slot_reads;    /* reads per batch; default 32768 */
  uint64_t slot_bases;    /* bases per batch; default 64 Mi */
  uint64_t hbm_budget;    /* bytes for the resident index plus all slots (0 = none) */
  uint32_t flags;         /* ZG_QUERY_NO_RECOVERY applies to every request */
  uint32_t reserved;
} zg_server_config;

typedef struct zg_request_stats {
  uint64_t ticket;
  uint64_t batch_id;        /* batches are numbered from 1 in admission order */
  uint32_t batch_requests;  /* requests in that batch */
  int32_t status;           /* ZG_OK or the ZG_ERR_* code of the batch */
  uint64_t n_reads;
  uint64_t seeds;           /* seed occurrences over the request's reads */
  double queued_ms;         /* submit -> admitted into a batch */
  double service_ms;        /* admitted -> results written */
  double latency_ms;        /* submit -> results written */
} zg_request_stats;

typedef struct zg_server_stats {
  uint64_t requests;        /* completed */
  uint64_t batches;
  uint64_t reads;
  uint64_t seeds;
  uint32_t max_batch_requests;
  uint32_t slots;
  double lookup_ms;         /* summed HIP-event time of the batches' lookups */
  uint64_t device_bytes;    /* resident index + slot buffers */
  uint64_t sweep_batches;   /* batches run as the sorted sweep (the index tuned
                               to sweep, k-mer seeds, banks resident, and the
                               batch large enough that the sweep is faster by
                               the server's model of both orders) */
} zg_server_stats;

void zg_server_config_default(zg_server_config *cfg);
/* The index must be loaded (zg_index_load / zg_index_load_ex) and stay open
 * and loaded until zg_server_destroy returns. Fails with ZG_ERR_BUDGET when
 * the resident part of the index and the slots do not fit hbm_budget. An
 * index loaded with ZG_MEMORY_SUBINDEX under a budget keeps the banks that do
 * not fit on the host: each batch streams them through HBM one at a time
 * (query_stream.rs:228-265), one batch at a time, while the other slots copy
 * and seed their next batches. */
int zg_server_create(zg_index *idx, const zg_server_config *cfg, zg_server **out);
/* Queues one request: reads seq[read_off[i] .. read_off[i+1]) (ASCII). The
 * input and the outputs stay owned by the caller and must remain valid until
 * the request is waited for. Outputs, each optional: totals[n_reads] (seed
 * occurrences per read), counts[n_reads x colours] (u32 matches per read and
 * colour, row-major) and aggregate[colours] (u64 sums over the reads). A
 * request larger than one slot fails with ZG_ERR_BUDGET. Never blocks. */
int zg_server_submit(zg_server *srv, const uint8_t *seq, const uint64_t *read_off,
                     uint64_t n_reads, uint64_t *totals, uint32_t *counts,
                     uint64_t *aggregate, uint64_t *ticket);
/* Blocks until the request is answered; returns its status. Each ticket is
 * waited for (or polled to completion) once. */
int zg_server_wait(zg_server *srv, uint64_t ticket, zg_request_stats *st);
/* 1 when answered (st filled, ticket released), 0 while queued or running. */
int zg_server_poll(zg_server *srv, uint64_t ticket, zg_request_stats *st);
int zg_server_get_stats(zg_server *srv, zg_server_stats *out);
/* Answers every queued request, then releases the slots. */
int zg_server_destroy(zg_server *srv);

/* Device characterisation (no reference counterpart): reads n_rows whole rows
 * of row_bytes (512, 1024, 2048, 4096 or 8192) at uniformly random row
 * indices of resident bank `bank`, four rows per wave in flight and no other
 * work, and returns the rate in GB/s: the ceiling of random row gathers at
 * that row size, which the lookup's own rate is compared with. */
int zg_gather_probe(zg_index *idx, uint32_t bank, uint32_t row_bytes,
                    uint64_t n_rows, double *gbs);
/* The same for the lookup's own row pattern: the H rows (H consecutive
 * segments) of n_keys random keys under bank `bank`'s PositionHash geometry,
 * one key per wave in flight, whole rows; banks of 512, 1024 or 2048-byte
 * rows. GB/s of row bytes. */
int zg_gather_probe_keys(zg_index *idx, uint32_t bank, uint64_t n_keys,
                         double *gbs);

/* Page-locked host memory for inputs and outputs of requests and batches
 * (DMA without staging copies). */
int zg_host_alloc(uint64_t bytes, void **out);
int zg_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
