// zg_format.h -- the reference on-disk index format: manifest.json, bank files
// (.zorbank.zst) and the abandoned-key recovery map (abandoned.zormap.zst).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "zg_json.h"

namespace zg {

// Error carrying one of the ZG_ERR_* codes of zindex_gpu.h.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string &msg);
inline void ensure(bool ok, const char *msg, int code = -3) {
  if (!ok) fail(code, msg);
}

constexpr char kPlaneMagic[9] = "ZORPOW3";   // + NUL (8 bytes)
constexpr char kLegacyMagic[9] = "ZORPOW1";
constexpr char kAbandonedMagic[9] = "ZORABN1";
constexpr char kAbandonedDeltaMagic[9] = "ZORABN2";
constexpr const char *kAbandonedFile = "abandoned.zormap.zst";
constexpr uint64_t kFingerprintSeed = 0x8E2D4F6A9B13C5D7ULL;
constexpr int kMaxHashes = 32;

// storage.rs:57-63
struct Color {
  std::string name, path;
  uint64_t occurrences = 0;
  uint64_t unique_kmers = 0;
};

// storage.rs:65-73
struct GroupMeta {
  std::vector<uint32_t> members;
  uint64_t width = 0, offset = 0, union_cardinality = 0, fallback_keys = 0;
  bool has_abandoned_keys = false;
  uint64_t abandoned_keys = 0;
};

// storage.rs:75-131
struct BankMeta {
  uint64_t component = 0;
  std::string file;
  uint64_t slots = 0, row_bits = 0;
  std::vector<GroupMeta> groups;
  bool fingerprint_first = false;
  bool routed = false;
  uint64_t routing_source_slots = 0, partition = 0, partitions = 1;

  uint64_t source_slots() const { return routed ? routing_source_slots : slots; }
  std::pair<uint64_t, uint64_t> source_group() const {
    return {component, source_slots()};
  }
  uint64_t row_bytes() const { return (row_bits + 7) / 8; }
  uint64_t payload_bytes() const { return slots * row_bytes(); }
  std::string expected_filename() const;
  static std::string filename(uint64_t slots, bool routed, uint64_t source_slots,
                              uint64_t partition, uint64_t partitions);
};

// storage.rs:133-258 (fields the query path reads; the rest is kept in `json`
// so that info and rewrites preserve them).
struct Manifest {
  uint32_t version = 0;
  bool abandoned_recovery = true;
  bool has_abandoned_file = false;
  std::string abandoned_file;
  uint64_t abandoned_hashes = 0, abandoned_origins = 0;
  std::string backend;
  uint64_t k = 0, hashes = 0;
  bool has_segment_length = false;
  uint64_t segment_length = 0;
  uint64_t seed = 0;
  uint64_t fingerprint_bits = 0;
  uint64_t index_mode = 0;
  uint64_t minimizer_size = 0;
  uint32_t minimizer_scheme = 0;  // 0 simd, 1 open-closed-mod
  bool homopolymer_compression = false;
  bool has_seed_pattern = false, has_randstrobemers = false;
  std::string seed_pattern;       // as written
  uint32_t seed_width = 0;        // parsed pattern (SpacedSeed)
  uint64_t seed_mask = 0;
  uint64_t strobe_min = 0, strobe_max = 0;  // RandstrobeConfig
  std::string randstrobe_scheme;
  uint64_t modimizer_sampling = 0, findere_z = 0;
  double slot_scale = 1.0;
  bool has_max_subindexes = false;
  uint64_t max_subindexes = 0;
  bool has_max_subindex_size = false;
  uint64_t max_subindex_size = 0;
  std::vector<Color> colors;
  std::vector<BankMeta> banks;
  Json json;  // the parsed document

  size_t classes() const;
  std::string merge_planner_label() const;
};

// Parses manifest.json text (serde rules: required fields, types, defaults).
// SpacedSeed::from_str (src/anchors.rs:15-43): a symmetric '#'/'_' pattern
// of 1..=31 bases with at least one '#'. mask holds 3 in the two bits of each
// retained base, the pattern's first base most significant. Returns "" or the
// reference's error message.
std::string parse_spaced_seed(const std::string &pattern, uint32_t *width, uint64_t *mask);

Manifest parse_manifest(const std::string &text);
// Index::load_mode validation (storage.rs:2301-2637), metadata only.
void validate_manifest(const Manifest &m);
Manifest load_manifest(const std::string &dir);
// Serializes a manifest built by this library.
std::string write_manifest_json(const Manifest &m, const Json &extra);

// ---- files ------------------------------------------------------------------
std::string read_file(const std::string &path);
void write_file(const std::string &path, const std::string &bytes);
std::string path_join(const std::string &dir, const std::string &name);

// zstd helpers
std::string zstd_decompress_all(const std::string &compressed);  // any frames
std::string zstd_decompress_all(const void *data, size_t n);
std::string zstd_compress(const void *data, size_t n, int level, bool checksum);

// Bank payload: load_modern_bank (storage.rs:2232-2256) with
// bank_io::read_parallel (bank_io.rs:122-291). Returns the decoded payload
// (slots * row_bytes bytes) after header validation.
std::vector<uint8_t> load_bank_payload(const std::string &dir, const BankMeta &meta,
                                       int threads);
// A bank file opened for streamed decoding (bank_io::read_parallel,
// bank_io.rs:235-291): the compressed file and its seek-table frames, whose
// raw offsets count the 40-byte header. No frames: a single-stream bank,
// decoded whole by load_bank_payload.
constexpr size_t kBankHeaderBytes = 40;
struct BankFrames {
  struct Frame {
    uint64_t compressed, compressed_len, raw, raw_len;
  };
  const uint8_t *data = nullptr;  // the bank file, memory-mapped read-only
  size_t size = 0;
  std::vector<Frame> frames;
  BankFrames() = default;
  BankFrames(const BankFrames &) = delete;
  BankFrames &operator=(const BankFrames &) = delete;
  BankFrames(BankFrames &&o) noexcept : data(o.data), size(o.size), frames(std::move(o.frames)) {
    o.data = nullptr;
    o.size = 0;
  }
  ~BankFrames();
};
BankFrames open_bank_frames(const std::string &dir, const BankMeta &meta);
// Decodes frame i (raw_len bytes) into dst; frame 0 starts with the header.
void decode_bank_frame(const BankFrames &b, size_t i, uint8_t *dst);
// The decoded 40-byte header against the manifest (storage.rs:2184-2231).
void check_bank_header(const uint8_t *header, const BankMeta &meta);
// Recovery map: sorted hashes, origin offsets and origins (dataset ids).
struct RecoveryMap {
  std::vector<uint32_t> hashes;
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> origins;
  size_t size() const { return hashes.size(); }
};

// A version-1 bank (storage.rs:2555-2612): the decoded stream holds the
// legacy header and payload, then for each group its fallback list (u64
// count == fallback_keys, then count x (u64 key, u64 value), keys strictly
// ascending, value below 2^width). Returns the payload; fallback[g] gets
// group g's (key, value) pairs (the caller checks fingerprints).
std::vector<uint8_t> load_v1_bank(const std::string &dir, const BankMeta &meta,
                                  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> &fallback);
// One recovered key of a version-1 group: (abandoned hash, global group id,
// member mask).
struct FallbackRecord {
  uint32_t hash;
  uint32_t group;
  uint64_t mask;
};
// visit_abandoned (storage.rs:1882-1923): records sorted by hash, origins =
// the members their masks select, sorted and deduplicated, per hash.
RecoveryMap recovery_from_records(std::vector<FallbackRecord> &records,
                                  const std::vector<std::vector<uint32_t>> &group_members);
// bank_io::Writer (bank_io.rs:34-120): 4 MiB checksummed frames + seek table.
void write_bank_file(const std::string &path, const BankMeta &meta,
                     const uint8_t *payload, uint64_t bytes, int threads);

// AbandonedReader (abandoned_io.rs:43-129) over the whole file.
RecoveryMap read_recovery_map(const std::string &dir, const Manifest &m);
// write_abandoned_map (storage.rs:2053-2094), ZORABN2.
void write_recovery_map(const std::string &path, const RecoveryMap &map);

}  // namespace zg
