// zg_format.cpp -- manifest parsing/validation and bank / recovery-map I/O in
// the reference on-disk format.
#include "zg_format.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <fstream>
#include <set>
#include <thread>

#include "zg_zstd.h"

namespace zg {

void fail(int code, const std::string &msg) { throw Error(code, msg); }

static bool is_pow2(uint64_t v) { return v && !(v & (v - 1)); }
static uint64_t ilog2(uint64_t v) {
  uint64_t r = 0;
  while (v >>= 1) r++;
  return r;
}

std::string BankMeta::filename(uint64_t slots, bool routed, uint64_t source_slots,
                               uint64_t partition, uint64_t partitions) {
  if (routed)
    return "bank-" + std::to_string(ilog2(source_slots)) + "-part-" +
           std::to_string(partition) + "-of-" + std::to_string(partitions) +
           ".zorbank.zst";
  return "bank-" + std::to_string(ilog2(slots)) + ".zorbank.zst";
}

std::string BankMeta::expected_filename() const {
  std::string name = filename(slots, routed, routing_source_slots, partition, partitions);
  if (component == 0) return name;
  return "component-" + std::to_string(component) + "-" + name;
}

size_t Manifest::classes() const {
  std::set<std::pair<uint64_t, uint64_t>> s;
  for (auto &b : banks) s.insert(b.source_group());
  return s.size();
}

std::string Manifest::merge_planner_label() const {
  const Json *p = json.get("merge_planner");
  if (!p || p->kind != Json::String) return "full";
  if (p->text == "all-pairs" || p->text == "full") return "full";
  return p->text;
}

// ---- typed field access (serde semantics) -----------------------------------
namespace {

[[noreturn]] void field_error(const std::string &name, const char *what) {
  fail(-3, "invalid manifest field `" + name + "`: " + what);
}

const Json &required(const Json &obj, const std::string &name) {
  const Json *v = obj.get(name);
  if (!v) fail(-3, "invalid manifest: missing field `" + name + "`");
  return *v;
}

uint64_t as_u64(const Json &v, const std::string &name) {
  if (v.kind != Json::Number) field_error(name, "expected an unsigned integer");
  const std::string &t = v.text;
  if (t.empty() || t[0] == '-' || t.find_first_of(".eE+") != std::string::npos)
    field_error(name, "expected an unsigned integer");
  errno = 0;
  char *end = nullptr;
  unsigned long long x = strtoull(t.c_str(), &end, 10);
  if (errno == ERANGE || *end) field_error(name, "integer out of range");
  return x;
}

double as_f64(const Json &v, const std::string &name) {
  if (v.kind != Json::Number) field_error(name, "expected a number");
  return strtod(v.text.c_str(), nullptr);
}

bool as_bool(const Json &v, const std::string &name) {
  if (v.kind != Json::Bool) field_error(name, "expected a boolean");
  return v.b;
}

std::string as_str(const Json &v, const std::string &name) {
  if (v.kind != Json::String) field_error(name, "expected a string");
  return v.text;
}

uint64_t req_u64(const Json &o, const std::string &n) { return as_u64(required(o, n), n); }
uint64_t opt_u64(const Json &o, const std::string &n, uint64_t dflt, bool *present = nullptr) {
  const Json *v = o.get(n);
  if (present) *present = v && !v->is_null();
  if (!v || v->is_null()) return dflt;
  return as_u64(*v, n);
}
bool opt_bool(const Json &o, const std::string &n, bool dflt) {
  const Json *v = o.get(n);
  if (!v) return dflt;
  return as_bool(*v, n);
}

}  // namespace

std::string parse_spaced_seed(const std::string &pattern, uint32_t *width, uint64_t *mask) {
  if (pattern.empty() || pattern.size() > 31) return "seed pattern width must be in 1..=31";
  for (char c : pattern)
    if (c != '#' && c != '_') return "seed pattern accepts only '#' and '_'";
  if (pattern.find('#') == std::string::npos) return "seed pattern must contain at least one '#'";
  if (!std::equal(pattern.begin(), pattern.end(), pattern.rbegin()))
    return "seed pattern must be symmetric under reversal for reverse-complement indexing";
  uint64_t mk = 0;
  for (char c : pattern) mk = (mk << 2) | (c == '#' ? 3u : 0u);
  *width = (uint32_t)pattern.size();
  *mask = mk;
  return "";
}

Manifest parse_manifest(const std::string &text) {
  Manifest m;
  try {
    m.json = JsonParser(text).parse();
  } catch (const JsonError &e) {
    fail(-3, e.what());
  }
  const Json &o = m.json;
  if (o.kind != Json::Object) fail(-3, "invalid manifest: expected an object");
  uint64_t version = req_u64(o, "version");
  if (version > UINT32_MAX) field_error("version", "out of range");
  m.version = (uint32_t)version;
  m.abandoned_recovery = opt_bool(o, "abandoned_recovery", true);
  if (const Json *f = o.get("abandoned_file"); f && !f->is_null()) {
    m.has_abandoned_file = true;
    m.abandoned_file = as_str(*f, "abandoned_file");
  }
  m.abandoned_hashes = opt_u64(o, "abandoned_hashes", 0);
  m.abandoned_origins = opt_u64(o, "abandoned_origins", 0);
  m.backend = as_str(required(o, "backend"), "backend");
  if (m.backend != "zor") field_error("backend", "unknown variant");
  m.k = req_u64(o, "k");
  m.hashes = req_u64(o, "hashes");
  m.segment_length = opt_u64(o, "segment_length", 0, &m.has_segment_length);
  m.seed = req_u64(o, "seed");
  m.fingerprint_bits = req_u64(o, "fingerprint_bits");
  m.index_mode = req_u64(o, "index_mode");
  if (m.index_mode > 255) field_error("index_mode", "out of range for u8");
  m.minimizer_size = req_u64(o, "minimizer_size");
  if (const Json *s = o.get("minimizer_scheme")) {
    std::string v = as_str(*s, "minimizer_scheme");
    if (v == "simd") m.minimizer_scheme = 0;
    else if (v == "open-closed-mod") m.minimizer_scheme = 1;
    else field_error("minimizer_scheme", "unknown variant");
  }
  m.homopolymer_compression = opt_bool(o, "homopolymer_compression", false);
  if (const Json *s = o.get("seed_pattern"); s && !s->is_null()) {
    m.seed_pattern = as_str(*s, "seed_pattern");
    m.has_seed_pattern = true;
    std::string err = parse_spaced_seed(m.seed_pattern, &m.seed_width, &m.seed_mask);
    if (!err.empty()) fail(-3, err);
  }
  if (const Json *s = o.get("randstrobemers"); s && !s->is_null()) {
    if (s->kind != Json::Object) field_error("randstrobemers", "expected an object");
    m.has_randstrobemers = true;
    m.strobe_min = req_u64(*s, "min_distance");
    m.strobe_max = req_u64(*s, "max_distance");
  }
  if (const Json *s = o.get("randstrobe_scheme"); s && !s->is_null())
    m.randstrobe_scheme = as_str(*s, "randstrobe_scheme");
  m.modimizer_sampling = req_u64(o, "modimizer_sampling");
  m.findere_z = req_u64(o, "findere_z");
  req_u64(o, "sketch_size");
  as_str(required(o, "sketch_method"), "sketch_method");
  if (const Json *p = o.get("merge_planner")) {
    std::string v = as_str(*p, "merge_planner");
    if (v != "full" && v != "all-pairs" && v != "streaming" && v != "naive")
      field_error("merge_planner", "unknown variant");
  }
  if (const Json *p = o.get("column_order")) {
    std::string v = as_str(*p, "column_order");
    if (v != "width" && v != "compression") field_error("column_order", "unknown variant");
  }
  m.slot_scale = as_f64(required(o, "slot_scale"), "slot_scale");
  m.max_subindexes = opt_u64(o, "max_subindexes", 0, &m.has_max_subindexes);
  m.max_subindex_size = opt_u64(o, "max_subindex_size", 0, &m.has_max_subindex_size);
  for (const char *n : {"singleton_logical_bits", "predicted_logical_bits",
                        "planned_groups", "rejected_groups", "exact_group_verifications"})
    req_u64(o, n);
  for (const char *n : {"sketch_seconds", "plan_seconds", "build_seconds"})
    as_f64(required(o, n), n);

  const Json &colors = required(o, "colors");
  if (colors.kind != Json::Array) field_error("colors", "expected an array");
  for (const Json &c : colors.items) {
    if (c.kind != Json::Object) field_error("colors", "expected objects");
    Color col;
    col.name = as_str(required(c, "name"), "name");
    col.path = as_str(required(c, "path"), "path");
    col.occurrences = req_u64(c, "occurrences");
    col.unique_kmers = req_u64(c, "unique_kmers");
    m.colors.push_back(std::move(col));
  }
  const Json &banks = required(o, "banks");
  if (banks.kind != Json::Array) field_error("banks", "expected an array");
  for (const Json &b : banks.items) {
    if (b.kind != Json::Object) field_error("banks", "expected objects");
    BankMeta bm;
    bm.component = opt_u64(b, "component", 0);
    bm.file = as_str(required(b, "file"), "file");
    bm.slots = req_u64(b, "slots");
    bm.row_bits = req_u64(b, "row_bits");
    bm.fingerprint_first = opt_bool(b, "fingerprint_first", false);
    if (const Json *r = b.get("routing"); r && !r->is_null()) {
      if (r->kind != Json::Object) field_error("routing", "expected an object");
      bm.routed = true;
      bm.routing_source_slots = req_u64(*r, "source_slots");
      bm.partition = req_u64(*r, "partition");
      bm.partitions = req_u64(*r, "partitions");
    }
    const Json &groups = required(b, "groups");
    if (groups.kind != Json::Array) field_error("groups", "expected an array");
    for (const Json &g : groups.items) {
      if (g.kind != Json::Object) field_error("groups", "expected objects");
      GroupMeta gm;
      const Json &mem = required(g, "members");
      if (mem.kind != Json::Array) field_error("members", "expected an array");
      for (const Json &x : mem.items) {
        uint64_t id = as_u64(x, "members");
        if (id > UINT32_MAX) fail(-3, "invalid/duplicate dataset mapping");
        gm.members.push_back((uint32_t)id);
      }
      gm.width = req_u64(g, "width");
      gm.offset = req_u64(g, "offset");
      gm.union_cardinality = req_u64(g, "union_cardinality");
      gm.fallback_keys = req_u64(g, "fallback_keys");
      gm.abandoned_keys = opt_u64(g, "abandoned_keys", 0, &gm.has_abandoned_keys);
      bm.groups.push_back(std::move(gm));
    }
    m.banks.push_back(std::move(bm));
  }
  return m;
}

void validate_manifest(const Manifest &m) {
  // storage.rs:2301-2304
  ensure(m.version >= 1 && m.version <= 10 && m.backend == "zor",
         "unsupported power-of-two index format");
  ensure(m.hashes >= 2 && m.hashes <= (uint64_t)kMaxHashes, "invalid hash count");
  bool seg_ok;
  if (m.version >= 9)
    seg_ok = !m.has_segment_length ||
             (is_pow2(m.segment_length) && m.segment_length <= 4096);
  else if (m.version == 8)
    seg_ok = m.has_segment_length && is_pow2(m.segment_length) &&
             m.segment_length <= 4096;
  else
    seg_ok = !m.has_segment_length;
  ensure(seg_ok, "invalid/missing segment length for index version");
  uint64_t b = m.fingerprint_bits;
  ensure(b == 1 || b == 2 || b == 4 || b == 8 || b == 16 || b == 32,
         "invalid fingerprint width");
  ensure(m.colors.size() <= UINT32_MAX, "dataset IDs must fit in 32 bits");
  if (m.version >= 2) {
    ensure(m.abandoned_recovery == m.has_abandoned_file,
           "inconsistent abandoned recovery metadata");
    ensure(!m.has_abandoned_file || m.abandoned_file == kAbandonedFile,
           "invalid abandoned map filename");
    ensure(m.abandoned_recovery ||
               (m.abandoned_hashes == 0 && m.abandoned_origins == 0),
           "disabled recovery has abandoned records");
  }
  // Manifest::features (storage.rs:243-258)
  ensure(!m.has_randstrobemers ||
             (m.version >= 6 && m.randstrobe_scheme == "strobealign-popcount8-symmetric64-downstream-v1"),
         "unsupported legacy/custom randstrobemer extraction; rebuild this index "
         "with the current downstream scheme");
  ensure(m.index_mode <= 3, "unsupported index mode");
  if ((m.has_seed_pattern || m.has_randstrobemers))
    ensure(m.index_mode == 0 || m.index_mode == 1,
           "spaced seeds and randstrobemers require --index-mode kmers or minimizers");
  // FeatureConfig::validate_anchors (src/lib.rs:177-205)
  if (m.has_seed_pattern) {
    const uint64_t length = m.index_mode == 1 ? m.minimizer_size : m.k;
    if (m.seed_width != length)
      fail(-3, "seed pattern width " + std::to_string(m.seed_width) + " must match " +
                              (m.index_mode == 1 ? "minimizer size m" : "k") + " (" +
                              std::to_string(length) + ")");
  }
  if (m.has_randstrobemers) {
    ensure(m.strobe_min > 0, "randstrobemer minimum distance must be greater than zero");
    ensure(m.strobe_min <= m.strobe_max,
           "randstrobemer minimum distance must not exceed maximum distance");
  }
  switch (m.index_mode) {
    case 3:
      ensure(m.findere_z < m.k, "invalid findere z");
      break;
    case 0:
    case 2:
      ensure(m.k > 0, "invalid k");
      break;
    case 1:
      ensure(m.k >= 1 && m.k <= 255 && m.k % 2 == 1 && m.minimizer_size <= m.k,
             "invalid minimizer window");
      break;
  }
  ensure(m.minimizer_size >= 1 && m.minimizer_size <= 31 && m.modimizer_sampling > 0,
         "invalid feature parameters");
  std::vector<int64_t> seen_bank(m.colors.size(), -1);
  std::vector<std::pair<uint64_t, uint64_t>> seen_class(m.colors.size());
  std::vector<std::pair<size_t, size_t>> groups_seen(m.colors.size(), {SIZE_MAX, 0});
  size_t classes = m.classes();
  ensure(!m.has_max_subindexes || (m.max_subindexes > 0 && classes <= m.max_subindexes),
         "invalid maximum subindex count");
  ensure(!m.has_max_subindex_size || is_pow2(m.max_subindex_size),
         "invalid maximum subindex size");
  std::pair<uint64_t, uint64_t> previous_group{0, 0};
  uint64_t next_partition = 0, partition_count = 1;
  size_t first_class_bank = 0;
  for (size_t bank_id = 0; bank_id < m.banks.size(); bank_id++) {
    const BankMeta &meta = m.banks[bank_id];
    ensure(meta.component == 0 || m.version >= 10, "unsupported index component");
    ensure(!meta.fingerprint_first || m.version >= 3, "unsupported bank layout");
    ensure(is_pow2(meta.slots) && meta.slots >= m.hashes && is_pow2(meta.source_slots()),
           "invalid bank size/order");
    uint64_t partition = meta.routed ? meta.partition : 0;
    uint64_t partitions = meta.routed ? meta.partitions : 1;
    ensure(!meta.routed ||
               (m.version >= 9 && partitions > 1 && is_pow2(partitions) &&
                meta.slots * partitions / partitions == meta.slots &&
                meta.slots * partitions == meta.source_slots()),
           "invalid hash partition geometry");
    if (meta.source_group() != previous_group) {
      ensure(meta.source_group() > previous_group &&
                 (previous_group == std::make_pair<uint64_t, uint64_t>(0, 0) ||
                  next_partition == partition_count),
             "invalid/incomplete partition order");
      previous_group = meta.source_group();
      next_partition = 0;
      partition_count = partitions;
      first_class_bank = bank_id;
    }
    ensure(partitions == partition_count && partition == next_partition &&
               partition < partitions,
           "invalid/duplicate/missing hash partition");
    next_partition++;
    if (m.has_segment_length)
      ensure(meta.slots % m.segment_length == 0 &&
                 meta.slots / m.segment_length >= m.hashes,
             "subindex too small for fixed segment geometry");
    ensure(meta.file == meta.expected_filename(), "invalid bank file name");
    uint64_t offset = 0;
    std::set<uint32_t> local_seen;
    for (size_t gid = 0; gid < meta.groups.size(); gid++) {
      const GroupMeta &g = meta.groups[gid];
      ensure(!g.members.empty() && g.members.size() <= 64 - b, "invalid membership width");
      uint64_t width = b + (g.members.size() == 1 ? 0 : g.members.size());
      ensure(g.offset == offset && g.width == width, "invalid group bit layout");
      offset += width;
      for (uint32_t id : g.members) {
        bool ok = id < m.colors.size() && local_seen.insert(id).second &&
                  (seen_bank[id] < 0 || seen_class[id] == meta.source_group());
        ensure(ok, "invalid/duplicate dataset mapping");
        if (partition == 0) {
          seen_bank[id] = (int64_t)bank_id;
          seen_class[id] = meta.source_group();
          groups_seen[id] = {bank_id, gid};
        } else {
          ensure(groups_seen[id].first != SIZE_MAX, "missing partition dataset");
          const GroupMeta &first =
              m.banks[groups_seen[id].first].groups[groups_seen[id].second];
          ensure(seen_class[id] == meta.source_group() && first.members == g.members,
                 "inconsistent partition dataset mapping");
        }
      }
      ensure(g.fallback_keys <= g.union_cardinality, "invalid fallback count");
      if (m.version >= 2) {
        ensure(g.has_abandoned_keys, "missing abandoned equation count");
        ensure(g.abandoned_keys <= g.union_cardinality, "invalid abandoned equation count");
        ensure(g.fallback_keys == (m.abandoned_recovery ? g.abandoned_keys : 0),
               "inconsistent abandoned equation count");
      }
    }
    ensure(offset == meta.row_bits && offset > 0, "invalid row width");
    if (partition > 0) {
      const BankMeta &first = m.banks[first_class_bank];
      size_t members = 0;
      for (auto &g : first.groups) members += g.members.size();
      ensure(first.row_bits == meta.row_bits && first.groups.size() == meta.groups.size() &&
                 local_seen.size() == members,
             "incomplete partition dataset mapping");
    }
    unsigned __int128 bytes = (unsigned __int128)meta.slots * meta.row_bytes();
    ensure(bytes <= (unsigned __int128)UINT64_MAX, "bank size overflow");
    ensure(!m.has_max_subindex_size || (uint64_t)bytes <= m.max_subindex_size,
           "subindex exceeds maximum uncompressed size");
  }
  bool all_seen = std::all_of(seen_bank.begin(), seen_bank.end(),
                              [](int64_t v) { return v >= 0; });
  ensure(!m.banks.empty() && all_seen && next_partition == partition_count,
         "incomplete dataset mapping");
}

// ---- files -------------------------------------------------------------------
std::string path_join(const std::string &dir, const std::string &name) {
  if (dir.empty()) return name;
  if (dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

std::string read_file(const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) fail(-2, "opening " + path + ": " + strerror(errno));
  std::string out;
  struct stat st;
  if (fstat(fileno(f), &st) == 0 && S_ISREG(st.st_mode)) out.reserve((size_t)st.st_size);
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  bool err = ferror(f);
  fclose(f);
  if (err) fail(-2, "reading " + path);
  return out;
}

void write_file(const std::string &path, const std::string &bytes) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) fail(-2, "creating " + path + ": " + strerror(errno));
  size_t n = fwrite(bytes.data(), 1, bytes.size(), f);
  bool ok = n == bytes.size() && fflush(f) == 0;
  if (fclose(f) != 0) ok = false;
  if (!ok) fail(-2, "writing " + path);
}

Manifest load_manifest(const std::string &dir) {
  Manifest m = parse_manifest(read_file(path_join(dir, "manifest.json")));
  validate_manifest(m);
  return m;
}

// ---- zstd ---------------------------------------------------------------------
std::string zstd_decompress_all(const std::string &compressed) {
  return zstd_decompress_all(compressed.data(), compressed.size());
}

std::string zstd_decompress_all(const void *data, size_t n) {
  ZSTD_DCtx *d = ZSTD_createDCtx();
  if (!d) fail(-5, "zstd: out of memory");
  std::string out;
  std::vector<char> buf(ZSTD_DStreamOutSize());
  ZSTD_inBuffer in{data, n, 0};
  // Decode until all input is consumed and the last call left output room
  // (nothing buffered); a final return of 0 means the last frame completed.
  size_t last = 1;
  bool flushed = false;
  while (!flushed) {
    ZSTD_outBuffer o{buf.data(), buf.size(), 0};
    size_t r = ZSTD_decompressStream(d, &o, &in);
    if (ZSTD_isError(r)) {
      std::string msg = ZSTD_getErrorName(r);
      ZSTD_freeDCtx(d);
      fail(-2, "zstd: " + msg);
    }
    out.append(buf.data(), o.pos);
    last = r;
    flushed = in.pos == in.size && o.pos < o.size;
  }
  ZSTD_freeDCtx(d);
  if (last != 0) fail(-2, "zstd: truncated stream");
  return out;
}

std::string zstd_compress(const void *data, size_t n, int level, bool checksum) {
  ZSTD_CCtx *c = ZSTD_createCCtx();
  if (!c) fail(-5, "zstd: out of memory");
  ZSTD_CCtx_setParameter(c, kZstdCompressionLevel, level);
  if (checksum) ZSTD_CCtx_setParameter(c, kZstdChecksumFlag, 1);
  std::string out(ZSTD_compressBound(n), '\0');
  size_t r = ZSTD_compress2(c, out.data(), out.size(), data, n);
  ZSTD_freeCCtx(c);
  if (ZSTD_isError(r)) fail(-2, std::string("zstd: ") + ZSTD_getErrorName(r));
  out.resize(r);
  return out;
}

// ---- bank container (bank_io.rs) ------------------------------------------------
namespace {
constexpr size_t kChunk = 4u << 20;
constexpr size_t kHeader = 40;
constexpr uint32_t kSkip = 0x184D2A50;
constexpr size_t kMaxTable = 64u << 20;

uint64_t le64(const uint8_t *p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}
uint32_t le32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}
void put64(std::string &s, uint64_t v) { s.append((const char *)&v, 8); }

uint64_t fnv1a(const uint8_t *p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ULL;
  for (size_t i = 0; i < n; i++) h = (h ^ p[i]) * 0x100000001b3ULL;
  return h;
}

struct Frame {
  uint64_t compressed, compressed_len, raw, raw_len;
};

// bank_io::table (bank_io.rs:122-197). Returns false when no table exists.
bool read_table(const uint8_t *p, uint64_t length, uint64_t raw_total, std::vector<Frame> &frames) {
  if (length < 16) return false;
  if (memcmp(p + length - 16, "ZORSEEK1", 8) != 0) return false;
  uint64_t len = le64(p + length - 8);
  ensure(len >= 40 && len <= kMaxTable && (len - 40) % 32 == 0 && len + 8 <= length,
         "invalid bank seek table length");
  uint64_t start = length - len - 8;
  ensure(le32(p + start) == kSkip && le32(p + start + 4) == len, "invalid skippable bank index");
  const uint8_t *t = p + start + 8;
  ensure(memcmp(t, "ZORFRM1\0", 8) == 0 && memcmp(t + len - 16, "ZORSEEK1", 8) == 0 &&
             le64(t + len - 8) == len,
         "invalid bank seek table tag");
  ensure(le64(t + len - 24) == fnv1a(t, len - 24), "bank seek table checksum mismatch");
  uint64_t count = le64(t + 8);
  ensure(count == (len - 40) / 32 && count != 0 && count <= raw_total, "invalid bank frame count");
  uint64_t compressed = 0, raw = 0;
  for (uint64_t i = 0; i < count; i++) {
    const uint8_t *v = t + 16 + 32 * i;
    Frame f{le64(v), le64(v + 8), le64(v + 16), le64(v + 24)};
    ensure(f.compressed == compressed && f.raw == raw && f.raw_len >= 1 &&
               f.raw_len <= kChunk && f.compressed_len > 0 &&
               f.compressed_len <= f.raw_len + f.raw_len / 128 + 1024,
           "invalid bank frame geometry");
    compressed += f.compressed_len;
    raw += f.raw_len;
    ensure(compressed <= start && raw <= raw_total, "bank frame exceeds payload");
    frames.push_back(f);
  }
  ensure(compressed == start && raw == raw_total && frames[0].raw_len >= kHeader,
         "incomplete bank frame coverage");
  return true;
}

void validate_header(const uint8_t *h, const BankMeta &meta, uint64_t expected) {
  ensure(memcmp(h, meta.fingerprint_first ? kPlaneMagic : kLegacyMagic, 8) == 0,
         "invalid bank magic");
  uint64_t want[4] = {meta.slots, meta.row_bits, (uint64_t)meta.groups.size(), expected};
  for (int i = 0; i < 4; i++)
    ensure(le64(h + 8 + 8 * i) == want[i], "bank header differs from manifest");
}

template <class F>
void parallel_for(size_t n, int threads, F &&f) {
  if (threads <= 1 || n <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  std::exception_ptr error;
  std::atomic<bool> failed{false};
  for (int t = 0; t < threads && (size_t)t < n; t++)
    pool.emplace_back([&] {
      try {
        for (size_t i; !failed && (i = next++) < n;) f(i);
      } catch (...) {
        if (!failed.exchange(true)) error = std::current_exception();
      }
    });
  for (auto &th : pool) th.join();
  if (error) std::rethrow_exception(error);
}
}  // namespace

std::vector<uint8_t> load_bank_payload(const std::string &dir, const BankMeta &meta,
                                       int threads) {
  const uint64_t expected = meta.payload_bytes();
  BankFrames bf = open_bank_frames(dir, meta);
  std::vector<uint8_t> data;
  if (!bf.frames.empty()) {
    data.assign(expected, 0);
    uint8_t header[kHeader];
    parallel_for(bf.frames.size(), threads, [&](size_t i) {
      const BankFrames::Frame &f = bf.frames[i];
      if (i == 0) {  // the raw stream opens with the header
        std::vector<uint8_t> tmp(f.raw_len);
        decode_bank_frame(bf, 0, tmp.data());
        memcpy(header, tmp.data(), kHeader);
        memcpy(data.data(), tmp.data() + kHeader, f.raw_len - kHeader);
      } else {
        decode_bank_frame(bf, i, data.data() + (f.raw - kHeader));
      }
    });
    validate_header(header, meta, expected);
    return data;
  }
  std::string raw = zstd_decompress_all(bf.data, bf.size);
  ensure(raw.size() >= kHeader, "truncated bank", -2);
  validate_header((const uint8_t *)raw.data(), meta, expected);
  ensure(raw.size() - kHeader >= expected, "truncated bank", -2);
  ensure(raw.size() - kHeader == expected, "unexpected bank data");
  data.assign(raw.begin() + kHeader, raw.end());
  return data;
}

BankFrames open_bank_frames(const std::string &dir, const BankMeta &meta) {
  BankFrames b;
  const std::string path = path_join(dir, meta.file);
  int fd = open(path.c_str(), O_RDONLY);
  if (fd < 0) fail(-2, "opening " + path + ": " + strerror(errno));
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    fail(-2, "reading " + path + ": " + strerror(errno));
  }
  b.size = (size_t)st.st_size;
  if (b.size) {
    void *p = mmap(nullptr, b.size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      close(fd);
      fail(-2, "mapping " + path + ": " + strerror(errno));
    }
    b.data = (const uint8_t *)p;
  }
  close(fd);
  std::vector<Frame> frames;
  if (read_table(b.data, b.size, meta.payload_bytes() + kHeader, frames) && frames.size() > 1)
    for (const Frame &f : frames) b.frames.push_back({f.compressed, f.compressed_len, f.raw, f.raw_len});
  return b;
}

BankFrames::~BankFrames() {
  if (data) munmap((void *)data, size);
}

void decode_bank_frame(const BankFrames &b, size_t i, uint8_t *dst) {
  const BankFrames::Frame &f = b.frames[i];
  const char *src = (const char *)b.data + f.compressed;
  ensure(ZSTD_findFrameCompressedSize(src, f.compressed_len) == f.compressed_len,
         "frame contains trailing data", -2);
  ensure(ZSTD_getFrameContentSize(src, f.compressed_len) == f.raw_len,
         "frame content size differs from seek table", -2);
  ZSTD_DCtx *d = ZSTD_createDCtx();
  size_t r = ZSTD_decompressDCtx(d, dst, f.raw_len, src, f.compressed_len);
  ZSTD_freeDCtx(d);
  ensure(!ZSTD_isError(r) && r == f.raw_len, "short bank frame", -2);
}

void check_bank_header(const uint8_t *header, const BankMeta &meta) {
  validate_header(header, meta, meta.payload_bytes());
}

std::vector<uint8_t> load_v1_bank(const std::string &dir, const BankMeta &meta,
                                  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> &fallback) {
  const uint64_t expected = meta.payload_bytes();
  BankFrames bf = open_bank_frames(dir, meta);
  std::string raw = zstd_decompress_all(bf.data, bf.size);
  ensure(raw.size() >= kHeader + expected, "truncated bank", -2);
  validate_header((const uint8_t *)raw.data(), meta, expected);
  const uint8_t *p = (const uint8_t *)raw.data() + kHeader + expected;
  const uint8_t *end = (const uint8_t *)raw.data() + raw.size();
  auto read_u64 = [&]() {
    ensure(end - p >= 8, "truncated bank", -2);
    uint64_t v = le64(p);
    p += 8;
    return v;
  };
  fallback.assign(meta.groups.size(), {});
  for (size_t g = 0; g < meta.groups.size(); g++) {
    const GroupMeta &gm = meta.groups[g];
    ensure(read_u64() == gm.fallback_keys, "invalid fallback length");
    for (uint64_t i = 0; i < gm.fallback_keys; i++) {
      const uint64_t key = read_u64(), value = read_u64();
      ensure(gm.width == 64 || value < (1ULL << gm.width), "invalid fallback payload width");
      ensure(fallback[g].empty() || fallback[g].back().first < key, "unsorted fallback keys");
      fallback[g].push_back({key, value});
    }
  }
  ensure(p == end, "unexpected bank data");
  return std::vector<uint8_t>(raw.begin() + kHeader, raw.begin() + kHeader + expected);
}

RecoveryMap recovery_from_records(std::vector<FallbackRecord> &records,
                                  const std::vector<std::vector<uint32_t>> &group_members) {
  std::sort(records.begin(), records.end(),
            [](const FallbackRecord &a, const FallbackRecord &b) { return a.hash < b.hash; });
  RecoveryMap map;
  std::vector<uint32_t> origins;
  for (size_t i = 0; i < records.size();) {
    const uint32_t h = records[i].hash;
    origins.clear();
    for (; i < records.size() && records[i].hash == h; i++) {
      const std::vector<uint32_t> &members = group_members[records[i].group];
      for (uint64_t mask = records[i].mask; mask; mask &= mask - 1) {
        const int bit = __builtin_ctzll(mask);
        ensure((size_t)bit < members.size(), "invalid abandoned membership bit");
        origins.push_back(members[bit]);
      }
    }
    std::sort(origins.begin(), origins.end());
    origins.erase(std::unique(origins.begin(), origins.end()), origins.end());
    ensure(!origins.empty(), "empty abandoned membership");
    map.hashes.push_back(h);
    map.origins.insert(map.origins.end(), origins.begin(), origins.end());
    map.off.push_back(map.origins.size());
  }
  return map;
}

void write_bank_file(const std::string &path, const BankMeta &meta, const uint8_t *payload,
                     uint64_t bytes, int threads) {
  // The raw stream is header + payload, cut into 4 MiB frames.
  std::string header(kPlaneMagic, 8);
  if (!meta.fingerprint_first) header.assign(kLegacyMagic, 8);
  put64(header, meta.slots);
  put64(header, meta.row_bits);
  put64(header, meta.groups.size());
  put64(header, bytes);
  uint64_t total = kHeader + bytes;
  size_t n_frames = (size_t)((total + kChunk - 1) / kChunk);
  std::vector<std::string> encoded(n_frames);
  parallel_for(n_frames, threads, [&](size_t i) {
    uint64_t lo = i * kChunk, hi = std::min<uint64_t>(total, lo + kChunk);
    std::string chunk;
    chunk.resize(hi - lo);
    for (uint64_t x = lo; x < hi;) {
      if (x < kHeader) {
        uint64_t n = std::min<uint64_t>(kHeader, hi) - x;
        memcpy(&chunk[x - lo], header.data() + x, n);
        x += n;
      } else {
        memcpy(&chunk[x - lo], payload + (x - kHeader), hi - x);
        x = hi;
      }
    }
    encoded[i] = zstd_compress(chunk.data(), chunk.size(), 0, true);
  });
  std::string out;
  std::string table("ZORFRM1\0", 8);
  put64(table, n_frames);
  uint64_t off = 0;
  for (size_t i = 0; i < n_frames; i++) {
    uint64_t lo = i * kChunk, hi = std::min<uint64_t>(total, lo + kChunk);
    put64(table, off);
    put64(table, encoded[i].size());
    put64(table, lo);
    put64(table, hi - lo);
    off += encoded[i].size();
  }
  put64(table, fnv1a((const uint8_t *)table.data(), table.size()));
  table.append("ZORSEEK1", 8);
  put64(table, table.size() + 8);
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) fail(-2, "creating " + path);
  bool ok = true;
  for (auto &e : encoded) ok &= fwrite(e.data(), 1, e.size(), f) == e.size();
  uint32_t skip[2] = {kSkip, (uint32_t)table.size()};
  ok &= fwrite(skip, 1, 8, f) == 8;
  ok &= fwrite(table.data(), 1, table.size(), f) == table.size();
  ok &= fflush(f) == 0;
  if (fclose(f) != 0) ok = false;
  if (!ok) fail(-2, "writing " + path);
}

// ---- recovery map --------------------------------------------------------------
namespace {
struct ByteReader {
  const uint8_t *p, *end;
  uint32_t u32() {
    ensure(end - p >= 4, "truncated abandoned map", -2);
    uint32_t v;
    memcpy(&v, p, 4);
    p += 4;
    return v;
  }
  // read_var_u32 (storage.rs:2097-2128)
  uint32_t var() {
    uint32_t value = 0;
    for (int shift = 0; shift <= 28; shift += 7) {
      ensure(p < end, "truncated abandoned map", -2);
      uint8_t byte = *p++;
      ensure(shift != 28 || byte <= 15, "abandoned variable integer overflow");
      value |= (uint32_t)(byte & 127) << shift;
      if (!(byte & 128)) return value;
    }
    fail(-3, "abandoned variable integer overflow");
  }
};
void put_var(std::string &out, uint32_t v) {
  while (v >= 128) {
    out += (char)((v & 127) | 128);
    v >>= 7;
  }
  out += (char)v;
}
}  // namespace

RecoveryMap read_recovery_map(const std::string &dir, const Manifest &m) {
  RecoveryMap map;
  ensure(m.has_abandoned_file && m.abandoned_file == kAbandonedFile,
         "invalid abandoned map filename");
  std::string file = read_file(path_join(dir, kAbandonedFile));
  std::string raw;
  if (file.size() >= 4 && (uint8_t)file[0] == 0x28 && (uint8_t)file[1] == 0xb5 &&
      (uint8_t)file[2] == 0x2f && (uint8_t)file[3] == 0xfd)
    raw = zstd_decompress_all(file);
  else
    raw = std::move(file);
  ensure(raw.size() >= 8, "invalid abandoned map magic");
  bool delta;
  if (memcmp(raw.data(), kAbandonedDeltaMagic, 8) == 0)
    delta = true;
  else if (memcmp(raw.data(), kAbandonedMagic, 8) == 0)
    delta = false;
  else
    fail(-3, "invalid abandoned map magic");
  ByteReader r{(const uint8_t *)raw.data() + 8, (const uint8_t *)raw.data() + raw.size()};
  uint64_t associations = 0;
  uint32_t previous = 0;
  map.hashes.reserve(m.abandoned_hashes);
  for (uint64_t rec = 0; rec < m.abandoned_hashes; rec++) {
    uint32_t hash;
    if (delta) {
      uint32_t d = r.var();
      ensure(rec == 0 || d != 0, "unsorted/duplicate abandoned hashes");
      ensure((uint64_t)previous + d <= UINT32_MAX, "abandoned hash delta overflow");
      hash = previous + d;
    } else {
      hash = r.u32();
    }
    ensure(rec == 0 || previous < hash, "unsorted/duplicate abandoned hashes");
    previous = hash;
    uint64_t count = delta ? r.var() : r.u32();
    ensure(count > 0 && count <= m.colors.size(), "invalid abandoned origin count");
    ensure(count <= (m.abandoned_origins > associations ? m.abandoned_origins - associations : 0),
           "abandoned origins exceed manifest");
    associations += count;
    uint32_t prev_origin = 0;
    size_t first = map.origins.size();
    for (uint64_t n = 0; n < count; n++) {
      uint32_t origin;
      if (delta) {
        uint32_t d = r.var();
        ensure(n == 0 || d != 0, "duplicate abandoned origins");
        ensure((uint64_t)prev_origin + d <= UINT32_MAX, "abandoned origin delta overflow");
        origin = prev_origin + d;
      } else {
        origin = r.u32();
      }
      ensure(origin < m.colors.size(), "invalid abandoned dataset ID");
      map.origins.push_back(origin);
      prev_origin = origin;
    }
    for (size_t i = first + 1; i < map.origins.size(); i++)
      ensure(map.origins[i - 1] < map.origins[i], "unsorted/duplicate abandoned origins");
    map.hashes.push_back(hash);
    map.off.push_back(map.origins.size());
  }
  ensure(associations == m.abandoned_origins, "abandoned origin count differs from manifest");
  ensure(r.p == r.end, "unexpected abandoned map data");
  return map;
}

void write_recovery_map(const std::string &path, const RecoveryMap &map) {
  std::string raw(kAbandonedDeltaMagic, 8);
  uint32_t previous = 0;
  for (size_t i = 0; i < map.size(); i++) {
    put_var(raw, map.hashes[i] - previous);
    previous = map.hashes[i];
    put_var(raw, (uint32_t)(map.off[i + 1] - map.off[i]));
    uint32_t prev_origin = 0;
    for (uint64_t j = map.off[i]; j < map.off[i + 1]; j++) {
      put_var(raw, map.origins[j] - prev_origin);
      prev_origin = map.origins[j];
    }
  }
  write_file(path, zstd_compress(raw.data(), raw.size(), 0, false));
}

std::string write_manifest_json(const Manifest &m, const Json &extra) {
  (void)m;
  std::string out;
  json_write(out, extra);
  return out;
}

}  // namespace zg
