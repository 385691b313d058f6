// zg_api.hip -- the C ABI of libzindex_amd.so: index handles, HBM residency,
// device-resident query batches and the CLI-equivalent commands.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <set>
#include <thread>

#include "zg_fastx.h"
#include "zg_hash.h"
#include "zg_internal.h"

namespace zg {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

int translate_exception() {
  try {
    throw;
  } catch (const Error &e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_error = "out of host memory";
    return ZG_ERR_NOMEM;
  } catch (const std::exception &e) {
    g_error = e.what();
    return ZG_ERR_INVALID;
  } catch (...) {
    g_error = "unknown error";
    return ZG_ERR_INVALID;
  }
}

void check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  int code = (e == hipErrorOutOfMemory) ? ZG_ERR_NOMEM : ZG_ERR_DEVICE;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t n) {
  reset();
  if (!n) n = 8;
  check_hip(hipMalloc(&p, n), "hipMalloc");
  bytes = n;
}

template <class T>
static void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty())
    check_hip(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
              "hipMemcpyAsync");
}

std::vector<uint64_t> ocm_inner_table(uint32_t s, uint64_t seed) {
  std::vector<uint64_t> t(256, 0);
  for (uint64_t x = 0; x < (1ULL << (2 * s)); x++)
    t[x] = mixsplit(canonical_encoded(x, (int)s), seed ^ kOcmInnerSeed);
  return t;
}

}  // namespace zg

// Host worker threads: the hardware count, capped by OMP_NUM_THREADS (or 16)
// so that a shared host is not oversubscribed.
unsigned zg_host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  unsigned cap = 16;
  if (const char *e = getenv("OMP_NUM_THREADS")) {
    long v = strtol(e, nullptr, 10);
    if (v > 0) cap = (unsigned)v;
  }
  n = n ? n : 1;
  return n < cap ? n : cap;
}

namespace zg {

static unsigned host_threads() { return zg_host_threads(); }

// Query support of this library for an index (the seed modes and schemes the GPU query runs).
static void check_query_support(const Manifest &m) {
  if (m.index_mode == 0 || m.index_mode == 2) {
    if (m.k > 32 && m.has_randstrobemers)
      fail(ZG_ERR_UNSUPPORTED, "randstrobemers of k-mers longer than 32 are not supported by the GPU query");
    if (m.index_mode == 2 && m.modimizer_sampling == 0)
      fail(ZG_ERR_FORMAT, "modimizer sampling must be positive");
  } else if (m.index_mode == 1) {
    if (m.minimizer_scheme != 1)
      fail(ZG_ERR_UNSUPPORTED, "only open-closed mod-minimizer indexes are supported by the GPU query");
  } else if (m.index_mode == 3) {
    if (m.k - m.findere_z > 32)
      fail(ZG_ERR_UNSUPPORTED, "findere s-mers longer than 32 are not supported by the GPU query");
    if (m.k > 255)
      fail(ZG_ERR_UNSUPPORTED, "findere windows longer than 255 are not supported by the GPU query");
  }
}

// get_bits / put_bits (storage.rs:3495-3538): little-endian bit fields of up
// to 64 bits at any bit offset of a row.
static inline uint64_t get_field(const uint8_t *row, uint64_t off, uint32_t width) {
  uint64_t v = 0;
  const uint32_t sh = (uint32_t)(off & 7);
  const uint32_t n = (sh + width + 7) / 8;
  const uint8_t *p = row + (off >> 3);
  for (uint32_t i = 0; i < n && i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
  v >>= sh;
  if (n == 9) v |= (uint64_t)p[8] << (64 - sh);
  return width == 64 ? v : v & ((1ULL << width) - 1);
}
static inline void put_field(uint8_t *row, uint64_t off, uint32_t width, uint64_t v) {
  for (uint32_t i = 0; i < width; i++)
    if ((v >> i) & 1) row[(off + i) >> 3] |= (uint8_t)(1u << ((off + i) & 7));
}

// The device keeps every bank in the plane layout with singleton columns
// first (the lookup's column model). A plane bank already in that order is
// uploaded as is; otherwise (legacy interleaved rows, fingerprint_first =
// false, where group g holds fp | mask << b at its offset, storage.rs:2864-2914,
// or columns not ordered singletons first) each row is rewritten into
// perm's column order. XOR is bitwise, so the XOR of rewritten rows is the
// rewrite of the XORed row and the lookup's results are unchanged.
static std::vector<uint8_t> plane_rows(const BankMeta &meta, const std::vector<uint8_t> &src,
                                       const std::vector<uint32_t> &perm, uint32_t fb) {
  const uint64_t rb = meta.row_bytes(), S = meta.slots;
  const uint32_t G = (uint32_t)meta.groups.size();
  std::vector<uint64_t> src_fp(G), src_mask(G), dst_mask(G);
  std::vector<uint32_t> mw(G);
  uint64_t acc = (uint64_t)G * fb;
  for (uint32_t g = 0; g < G; g++) {
    const GroupMeta &gm = meta.groups[g];
    mw[g] = gm.members.size() > 1 ? (uint32_t)gm.members.size() : 0;
    if (meta.fingerprint_first) {
      src_fp[g] = (uint64_t)g * fb;
      src_mask[g] = acc;
      acc += mw[g];
    } else {
      src_fp[g] = gm.offset;
      src_mask[g] = gm.offset + fb;
    }
  }
  acc = (uint64_t)G * fb;
  for (uint32_t c = 0; c < G; c++) {
    dst_mask[c] = acc;
    acc += mw[perm[c]];
  }
  std::vector<uint8_t> out(src.size(), 0);
  const uint64_t block = 4096;
  std::atomic<uint64_t> next{0};
  auto work = [&] {
    for (uint64_t b; (b = next++) * block < S;)
      for (uint64_t r = b * block; r < std::min(S, (b + 1) * block); r++) {
        const uint8_t *in = src.data() + r * rb;
        uint8_t *o = out.data() + r * rb;
        for (uint32_t c = 0; c < G; c++) {
          const uint32_t g = perm[c];
          put_field(o, (uint64_t)c * fb, fb, get_field(in, src_fp[g], fb));
          if (mw[g]) put_field(o, dst_mask[c], mw[g], get_field(in, src_mask[g], mw[g]));
        }
      }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < host_threads(); t++) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  return out;
}

constexpr uint64_t kStageBytes = 64ULL << 20;  // one page-locked upload chunk
constexpr uint64_t kBankPad = 4096;             // zeroed tail of every bank buffer

static void ensure_staging(zg_index *idx) {
  for (int k = 0; k < 2; k++) {
    if (!idx->staging[k])
      check_hip(hipHostMalloc(&idx->staging[k], kStageBytes, hipHostMallocDefault), "hipHostMalloc");
    if (!idx->staging_ev[k])
      check_hip(hipEventCreateWithFlags(&idx->staging_ev[k], hipEventDisableTiming), "hipEventCreate");
  }
}

template <class F>
static void parallel_ranges(uint64_t n, uint64_t grain, F &&f) {
  const uint64_t parts = std::max<uint64_t>(1, std::min<uint64_t>(host_threads(), n / grain));
  if (parts == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (uint64_t t = 0; t < parts; t++)
    pool.emplace_back([&, t] { f(n * t / parts, n * (t + 1) / parts); });
  for (auto &th : pool) th.join();
}

// Host bytes -> device through the two page-locked chunks: one chunk is
// filled (a parallel memcpy) while the other is in flight.
static void staged_copy(zg_index *idx, uint8_t *dst, const uint8_t *src, uint64_t n,
                        hipStream_t s) {
  ensure_staging(idx);
  int k = 0;
  for (uint64_t at = 0; at < n; at += kStageBytes, k ^= 1) {
    const uint64_t len = std::min(kStageBytes, n - at);
    check_hip(hipEventSynchronize(idx->staging_ev[k]), "staging");
    uint8_t *buf = (uint8_t *)idx->staging[k];
    parallel_ranges(len, 4ULL << 20, [&](uint64_t a, uint64_t b) { memcpy(buf + a, src + at + a, b - a); });
    check_hip(hipMemcpyAsync(dst + at, buf, len, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    check_hip(hipEventRecord(idx->staging_ev[k], s), "hipEventRecord");
  }
}

// A bank file -> device, frame by frame (bank_io::read_parallel,
// bank_io.rs:235-291): the frames of one chunk decode in parallel into a
// page-locked chunk, which is copied while the next chunk decodes.
static void stream_frames(zg_index *idx, const BankFrames &bf, const BankMeta &meta, uint8_t *dst,
                          hipStream_t s) {
  ensure_staging(idx);
  const auto &fr = bf.frames;
  int k = 0;
  for (size_t i = 0; i < fr.size(); k ^= 1) {
    const uint64_t r0 = fr[i].raw;
    size_t j = i;
    while (j < fr.size() && fr[j].raw + fr[j].raw_len - r0 <= kStageBytes) j++;
    ensure(j > i, "bank frame larger than an upload chunk", ZG_ERR_FORMAT);
    check_hip(hipEventSynchronize(idx->staging_ev[k]), "staging");
    uint8_t *buf = (uint8_t *)idx->staging[k];
    std::atomic<size_t> next{i};
    std::exception_ptr error;
    std::mutex mu;
    auto work = [&] {
      try {
        for (size_t f; (f = next++) < j;) decode_bank_frame(bf, f, buf + (fr[f].raw - r0));
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!error) error = std::current_exception();
      }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < std::min<size_t>(host_threads(), j - i); t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (error) std::rethrow_exception(error);
    const uint64_t r1 = fr[j - 1].raw + fr[j - 1].raw_len;
    uint64_t skip = 0;
    if (r0 == 0) {  // the raw stream opens with the bank header
      check_bank_header(buf, meta);
      skip = kBankHeaderBytes;
    }
    check_hip(hipMemcpyAsync(dst + (r0 + skip - kBankHeaderBytes), buf + skip, r1 - r0 - skip,
                             hipMemcpyHostToDevice, s),
              "hipMemcpyAsync");
    check_hip(hipEventRecord(idx->staging_ev[k], s), "hipEventRecord");
    i = j;
  }
}

// Device rows start on 16-byte boundaries, the widest lane load: a bank whose
// row length is not a multiple of 16 is stored with a padded pitch.
static uint64_t row_pitch(const BankMeta &m) { return (m.row_bytes() + 15) / 16 * 16; }
static uint64_t device_bytes(const BankMeta &m) { return m.slots * row_pitch(m); }

// Rows of rb bytes copied to a pitch of `pitch` bytes, the padding zeroed.
static std::vector<uint8_t> pitched_rows(const uint8_t *src, uint64_t slots, uint64_t rb,
                                         uint64_t pitch) {
  std::vector<uint8_t> out(slots * pitch);
  parallel_ranges(slots, 1 << 16, [&](uint64_t a, uint64_t b) {
    for (uint64_t r = a; r < b; r++) {
      memcpy(out.data() + r * pitch, src + r * rb, rb);
      memset(out.data() + r * pitch + rb, 0, pitch - rb);
    }
  });
  return out;
}

// Bank bi's payload in the device layout into dst (plus a zeroed pad), from
// the host copy when the index has one, else from the bank file.
static void put_bank(zg_index *idx, size_t bi, uint8_t *dst, hipStream_t s) {
  const BankMeta &meta = idx->m.banks[bi];
  const uint64_t bytes = meta.payload_bytes(), pitch = row_pitch(meta);
  const bool host = bi < idx->payload.size() && idx->payload[bi].size() == bytes;
  if (!idx->as_is[bi] || pitch != meta.row_bytes()) {
    // rows rewritten on the host: legacy layout and/or padded pitch
    std::vector<uint8_t> tmp;
    if (!host) tmp = load_bank_payload(idx->dir, meta, (int)host_threads());
    const std::vector<uint8_t> &src = host ? idx->payload[bi] : tmp;
    std::vector<uint8_t> rows;
    if (!idx->as_is[bi])
      rows = plane_rows(meta, src, idx->perm[bi], (uint32_t)idx->m.fingerprint_bits);
    const uint8_t *plane = idx->as_is[bi] ? src.data() : rows.data();
    if (pitch != meta.row_bytes()) {
      std::vector<uint8_t> padded = pitched_rows(plane, meta.slots, meta.row_bytes(), pitch);
      staged_copy(idx, dst, padded.data(), padded.size(), s);
    } else {
      staged_copy(idx, dst, plane, bytes, s);
    }
    check_hip(hipStreamSynchronize(s), "upload");  // host rows end here
  } else if (host) {
    staged_copy(idx, dst, idx->payload[bi].data(), bytes, s);
  } else {
    BankFrames bf = open_bank_frames(idx->dir, meta);
    if (bf.frames.empty()) {  // a single-stream bank decodes whole
      std::vector<uint8_t> tmp = load_bank_payload(idx->dir, meta, (int)host_threads());
      staged_copy(idx, dst, tmp.data(), bytes, s);
      check_hip(hipStreamSynchronize(s), "upload");
    } else {
      stream_frames(idx, bf, meta, dst, s);
    }
  }
  check_hip(hipMemsetAsync(dst + device_bytes(meta), 0, kBankPad, s), "hipMemsetAsync");
}

bool any_streamed(const zg_index *idx) {
  for (bool b : idx->streamed)
    if (b) return true;
  return false;
}

void make_resident(zg_index *idx, size_t i, hipStream_t s) {
  if (!idx->streamed[i] || idx->stream_holds == (int)i) return;
  idx->stream_holds = -1;
  put_bank(idx, i, idx->stream_data.as<uint8_t>(), s);
  idx->stream_holds = (int)i;
}

void upload_index(zg_index *idx, uint64_t budget, uint32_t memory) {
  const Manifest &m = idx->m;
  check_query_support(m);
  ensure(memory == ZG_MEMORY_ALL || memory == ZG_MEMORY_SUBINDEX, "invalid query memory mode",
         ZG_ERR_INVALID);
  idx->spaced = m.has_seed_pattern ? spaced_runs(m.seed_width, m.seed_mask) : SpacedRuns{};
  int dev = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  idx->device = dev;
  idx->cu_count = (uint32_t)prop.multiProcessorCount;
  if (!idx->stream) check_hip(hipStreamCreateWithFlags(&idx->stream, hipStreamNonBlocking), "hipStreamCreate");
  const uint32_t fb = (uint32_t)m.fingerprint_bits;
  const size_t nb = m.banks.size();
  // residency plan
  uint64_t payload_total = 0, largest = 0;
  for (auto &b : m.banks) {
    payload_total += device_bytes(b) + kBankPad;
    largest = std::max(largest, device_bytes(b) + kBankPad);
  }
  uint64_t rec_bytes = 0;
  if (idx->recover && idx->have_rec) rec_bytes = idx->rec.size() * 28 + idx->rec.origins.size() * 8;
  idx->streamed.assign(nb, false);
  if (budget && payload_total + rec_bytes > budget) {
    if (memory != ZG_MEMORY_SUBINDEX)
      fail(ZG_ERR_BUDGET, "index needs " + std::to_string(payload_total + rec_bytes) +
                              " bytes of HBM, budget is " + std::to_string(budget));
    if (rec_bytes + largest > budget)
      fail(ZG_ERR_BUDGET, "the recovery map and the largest subindex need " +
                              std::to_string(rec_bytes + largest) + " bytes of HBM, budget is " +
                              std::to_string(budget));
    uint64_t left = budget - rec_bytes - largest;  // beside one streaming buffer
    for (size_t bi = 0; bi < nb; bi++) {
      const uint64_t b = device_bytes(m.banks[bi]) + kBankPad;
      if (b <= left) left -= b;
      else idx->streamed[bi] = true;
    }
  }
  hipStream_t s = idx->stream;
  // the sweep's tables (and bit-plane copies) describe the banks uploaded here
  idx->sw_row_words = 0;
  idx->sw_classes.clear();
  idx->dbanks.clear();
  idx->dbanks.resize(nb);
  idx->perm.assign(nb, {});
  idx->as_is.assign(nb, true);
  idx->stream_holds = -1;
  idx->stream_data.reset();
  if (any_streamed(idx)) idx->stream_data.alloc(largest);
  const uint32_t n_colors = (uint32_t)m.colors.size();
  std::vector<std::vector<uint32_t>> bank_color_col(nb);  // device columns
  std::vector<std::vector<uint32_t>> fd_keys(nb);         // findere colour -> bit
  idx->fd_bank_member_bit.assign(nb, 0);
  for (size_t bi = 0; bi < nb; bi++) {
    const BankMeta &meta = m.banks[bi];
    DeviceBank &db = idx->dbanks[bi];
    const uint32_t G = (uint32_t)meta.groups.size();
    // device column order: singleton groups, then multi-member groups
    std::vector<uint32_t> &perm = idx->perm[bi];
    for (uint32_t g = 0; g < G; g++)
      if (meta.groups[g].members.size() == 1) perm.push_back(g);
    const uint32_t first_merged = (uint32_t)perm.size();
    for (uint32_t g = 0; g < G; g++)
      if (meta.groups[g].members.size() > 1) perm.push_back(g);
    bool as_is = true;
    for (uint32_t c = 0; c < G; c++) {
      const GroupMeta &gm = meta.groups[perm[c]];
      as_is &= perm[c] == c;
      // legacy rows of singletons only coincide with the plane layout
      if (!meta.fingerprint_first) as_is &= gm.members.size() == 1 && gm.offset == (uint64_t)c * fb;
    }
    idx->as_is[bi] = as_is;
    std::vector<uint32_t> col_color(G, kNone), mg_off{0}, mg_members;
    std::vector<uint64_t> mg_mask_bit;
    std::vector<uint32_t> color_col(n_colors, kNone), color_bit(n_colors, 0);
    uint64_t mask_bit = (uint64_t)G * fb;
    for (uint32_t c = 0; c < G; c++) {
      const GroupMeta &gm = meta.groups[perm[c]];
      for (size_t j = 0; j < gm.members.size(); j++) {
        color_col[gm.members[j]] = c;
        color_bit[gm.members[j]] = (uint32_t)j;
      }
      if (c < first_merged) {
        col_color[c] = gm.members[0];
      } else {
        mg_mask_bit.push_back(mask_bit);
        mask_bit += gm.members.size();
        for (uint32_t x : gm.members) mg_members.push_back(x);
        mg_off.push_back((uint32_t)mg_members.size());
      }
    }
    uint8_t *home;
    if (idx->streamed[bi]) {
      home = idx->stream_data.as<uint8_t>();
    } else {
      db.data.alloc(device_bytes(meta) + kBankPad);
      home = db.data.as<uint8_t>();
      put_bank(idx, bi, home, s);
    }
    if (m.index_mode == 3) {
      // hit-row bits: device column for singletons, then one bit per member
      const uint32_t tile_bits = 64 * fields_per_lane(fb);
      const uint32_t member_bit = ((G + tile_bits - 1) / tile_bits) * tile_bits;
      std::vector<uint32_t> key(n_colors + 1, kNone);
      for (uint32_t c = 0; c < G; c++) {
        const GroupMeta &gm = meta.groups[perm[c]];
        if (c < first_merged) key[gm.members[0]] = c;
        else
          for (size_t j = 0; j < gm.members.size(); j++)
            key[gm.members[j]] = member_bit + mg_off[c - first_merged] + (uint32_t)j;
      }
      key[n_colors] = member_bit + (uint32_t)mg_members.size();  // bits in use
      fd_keys[bi] = std::move(key);
      idx->fd_bank_member_bit[bi] = member_bit;
    }
    upload(db.col_color, col_color, s);
    upload(db.mg_off, mg_off, s);
    upload(db.mg_members, mg_members, s);
    upload(db.mg_mask_bit, mg_mask_bit, s);
    upload(db.color_col, color_col, s);
    bank_color_col[bi] = color_col;
    upload(db.color_bit, color_bit, s);
    BankDev &v = db.view;
    v.data = home;
    v.row_bytes = row_pitch(meta);
    v.plane_bytes = ((uint64_t)G * fb + 7) / 8;
    v.groups = G;
    v.first_merged = first_merged;
    v.hashes = (uint32_t)m.hashes;
    v.partition = (uint32_t)(meta.routed ? meta.partition : 0);
    v.partitions = (uint32_t)(meta.routed ? meta.partitions : 1);
    uint32_t tile_fields = 64 * fields_per_lane(fb);
    v.n_tiles = (G + tile_fields - 1) / tile_fields;
    v.geo = segment_geometry(meta.slots, (uint32_t)m.hashes,
                             m.has_segment_length ? m.segment_length : 0);
    v.col_color = db.col_color.as<uint32_t>();
    v.mg_off = db.mg_off.as<uint32_t>();
    v.mg_members = db.mg_members.as<uint32_t>();
    v.mg_mask_bit = db.mg_mask_bit.as<uint64_t>();
    v.color_col = db.color_col.as<uint32_t>();
    v.color_bit = db.color_bit.as<uint32_t>();
    // the host tables of this bank end with the loop
    check_hip(hipStreamSynchronize(s), "upload");
  }
  // Column map ids: banks whose colour -> column maps are equal (the
  // partitions of one size class) share an id.
  std::vector<std::vector<uint32_t>> maps;
  for (size_t bi = 0; bi < m.banks.size(); bi++) {
    size_t id = 0;
    while (id < maps.size() && maps[id] != bank_color_col[bi]) id++;
    if (id == maps.size()) maps.push_back(bank_color_col[bi]);
    idx->dbanks[bi].view.colmap = (uint32_t)id;
  }
  std::vector<std::vector<uint32_t>>().swap(bank_color_col);
  // findere regions: one per distinct colour -> bit map
  idx->fd_regions.clear();
  idx->fd_bank_region.assign(nb, 0);
  idx->fd_row_words = 0;
  if (m.index_mode == 3) {
    std::vector<size_t> first;  // a bank of each region
    std::vector<uint32_t> bit_color;
    for (size_t bi = 0; bi < nb; bi++) {
      size_t id = 0;
      while (id < first.size() && fd_keys[first[id]] != fd_keys[bi]) id++;
      if (id == first.size()) {
        zg_index::FdRegion rg;
        rg.bits = fd_keys[bi][n_colors];
        rg.off = (uint32_t)idx->fd_row_words;
        rg.color_at = bit_color.size();
        idx->fd_row_words += (rg.bits + 31) / 32;
        bit_color.resize(bit_color.size() + rg.bits, kNone);
        for (uint32_t c = 0; c < n_colors; c++)
          if (fd_keys[bi][c] != kNone) bit_color[rg.color_at + fd_keys[bi][c]] = c;
        idx->fd_regions.push_back(rg);
        first.push_back(bi);
      }
      idx->fd_bank_region[bi] = (uint32_t)id;
    }
    if (bit_color.empty()) bit_color.push_back(kNone);
    upload(idx->fd_bit_color, bit_color, s);
    check_hip(hipStreamSynchronize(s), "upload");
  }
  // recovery directory (the query-side OriginMap, abandoned.rs:29-206)
  idx->rec_view = RecDev{};
  if (idx->recover && idx->have_rec && idx->rec.size()) {
    const RecoveryMap &r = idx->rec;
    uint64_t n = r.size();
    uint64_t cap = 4;
    while (cap < (4 * n + 2) / 3) cap <<= 1;
    uint32_t bits = 0;
    while ((1ULL << bits) < cap) bits++;
    ensure(bits <= 32, "recovery directory too large", ZG_ERR_UNSUPPORTED);
    ensure(maps.size() <= 255, "more than 255 distinct bank column maps", ZG_ERR_UNSUPPORTED);
    // each colour's (map id, column); a colour sits in one size class
    std::vector<uint32_t> color_tag(n_colors, kNone);
    for (size_t id = 0; id < maps.size(); id++)
      for (uint32_t c = 0; c < n_colors; c++)
        if (maps[id][c] != kNone) {
          ensure(maps[id][c] < (1u << 24), "bank with 2^24 or more columns", ZG_ERR_UNSUPPORTED);
          color_tag[c] = ((uint32_t)id << 24) | maps[id][c];
        }
    std::vector<uint64_t> dir(cap, (uint64_t)kNone);
    std::vector<uint32_t> post, post_col, filter(kRecFilterWords, 0);
    const uint32_t probes = rec_filter_probes(n);
    post.reserve(n + r.origins.size());
    post_col.reserve(n + r.origins.size());
    for (uint64_t i = 0; i < n; i++) {
      uint32_t pi = (uint32_t)post.size();
      post.push_back((uint32_t)(r.off[i + 1] - r.off[i]));
      post_col.push_back(kNone);
      for (uint64_t j = r.off[i]; j < r.off[i + 1]; j++) {
        post.push_back(r.origins[j]);
        post_col.push_back(r.origins[j] < n_colors ? color_tag[r.origins[j]] : kNone);
      }
      uint64_t slot = r.hashes[i] >> (32 - bits);
      while ((uint32_t)dir[slot] != kNone) slot = (slot + 1) & (cap - 1);
      dir[slot] = ((uint64_t)r.hashes[i] << 32) | pi;
      for (uint32_t j = 0; j < probes; j++) {
        const uint32_t fb = rec_filter_bit(r.hashes[i], j);
        filter[fb / 32] |= 1u << (fb % 32);
      }
    }
    upload(idx->rec_dir, dir, s);
    upload(idx->rec_filter, filter, s);
    idx->rec_view.filter = idx->rec_filter.as<uint32_t>();
    idx->rec_view.filter_probes = probes;
    upload(idx->rec_post, post, s);
    upload(idx->rec_post_col, post_col, s);
    idx->rec_view.dir = idx->rec_dir.as<uint64_t>();
    idx->rec_view.post = idx->rec_post.as<uint32_t>();
    idx->rec_view.post_col = idx->rec_post_col.as<uint32_t>();
    idx->rec_view.cap_log2 = bits;
    idx->rec_view.enabled = 1;
  }
  if (m.index_mode == 1) {
    OcmParams P = ocm_parameters((int)m.k, (int)m.minimizer_size);
    upload(idx->inner_table, ocm_inner_table((uint32_t)P.s, m.seed), s);
  }
  check_hip(hipStreamSynchronize(s), "upload");
  idx->on_device = true;
}

uint64_t resident_bytes(const zg_index *idx) {
  uint64_t n = idx->rec_dir.bytes + idx->rec_post.bytes + idx->rec_post_col.bytes + idx->rec_filter.bytes +
               idx->inner_table.bytes + idx->stream_data.bytes;
  for (auto &d : idx->dbanks)
    n += d.data.bytes + d.col_color.bytes + d.mg_off.bytes + d.mg_members.bytes +
         d.mg_mask_bit.bytes + d.color_col.bytes + d.color_bit.bytes + d.planes.bytes;
  return n;
}

}  // namespace zg

using namespace zg;

zg_index::~zg_index() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (int k = 0; k < 2; k++) {
    if (staging_ev[k]) (void)hipEventDestroy(staging_ev[k]);
    if (staging[k]) (void)hipHostFree(staging[k]);
  }
  dbanks.clear();
  stream_data.reset();
  rec_dir.reset();
  rec_post.reset();
  rec_post_col.reset();
  rec_filter.reset();
  inner_table.reset();
  if (stream) (void)hipStreamDestroy(stream);
}

zg_batch::~zg_batch() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (sw_exec) (void)hipGraphExecDestroy(sw_exec);
  if (sw_graph) (void)hipGraphDestroy(sw_graph);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : sw_ev) (void)hipEventDestroy(e);
}

extern "C" const char *zg_last_error(void) { return g_error.c_str(); }
extern "C" const char *zg_version(void) { return "zindex_amd 0.1.0 (gfx950)"; }

extern "C" int zg_index_open(const char *dir, zg_index **out) {
  try {
    ensure(dir && out, "null argument", ZG_ERR_INVALID);
    auto idx = std::make_unique<zg_index>();
    idx->m = load_manifest(dir);
    idx->dir = dir;
    *out = idx.release();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_close(zg_index *idx) {
  delete idx;
  return 0;
}

extern "C" int zg_index_get_meta(const zg_index *idx, zg_index_meta *o) {
  try {
    ensure(idx && o, "null argument", ZG_ERR_INVALID);
    const Manifest &m = idx->m;
    memset(o, 0, sizeof *o);
    o->version = m.version;
    o->k = (uint32_t)m.k;
    o->hashes = (uint32_t)m.hashes;
    o->fingerprint_bits = (uint32_t)m.fingerprint_bits;
    o->index_mode = (uint32_t)m.index_mode;
    o->minimizer_size = (uint32_t)m.minimizer_size;
    o->minimizer_scheme = m.minimizer_scheme;
    o->n_colors = (uint32_t)m.colors.size();
    o->n_banks = (uint32_t)m.banks.size();
    o->n_classes = (uint32_t)m.classes();
    o->seed = m.seed;
    o->segment_length = m.has_segment_length ? m.segment_length : 0;
    o->abandoned_recovery = m.abandoned_recovery;
    o->abandoned_hashes = m.abandoned_hashes;
    o->abandoned_origins = m.abandoned_origins;
    for (auto &b : m.banks) o->payload_bytes += b.payload_bytes();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_color(const zg_index *idx, uint32_t color, const char **name,
                              uint64_t *unique_kmers) {
  try {
    ensure(idx && color < idx->m.colors.size(), "invalid colour", ZG_ERR_INVALID);
    if (name) *name = idx->m.colors[color].name.c_str();
    if (unique_kmers) *unique_kmers = idx->m.colors[color].unique_kmers;
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

namespace zg {
// Version 1 (storage.rs:2555-2612, 2665-2674): banks carry their groups'
// fallback keys after the payload, and the recovery map is built from them.
static void load_v1(zg_index *idx) {
  const Manifest &m = idx->m;
  const uint32_t b = (uint32_t)m.fingerprint_bits;
  const uint64_t fp_mask = (1ULL << b) - 1;
  std::vector<FallbackRecord> records;
  std::vector<std::vector<uint32_t>> members;
  idx->payload.assign(m.banks.size(), {});
  for (size_t bi = 0; bi < m.banks.size(); bi++) {
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> fallback;
    idx->payload[bi] = load_v1_bank(idx->dir, m.banks[bi], fallback);
    for (size_t g = 0; g < m.banks[bi].groups.size(); g++) {
      const GroupMeta &gm = m.banks[bi].groups[g];
      const uint32_t gid = (uint32_t)members.size();
      members.push_back(gm.members);
      for (auto &[key, value] : fallback[g]) {
        ensure((value & fp_mask) == fingerprint(key, m.seed, b), "invalid fallback fingerprint");
        const uint64_t mask = gm.members.size() == 1 ? 1 : value >> b;
        ensure(mask != 0, "empty fallback membership");
        records.push_back({abandoned_hash(key, m.seed), gid, mask});
      }
    }
  }
  idx->payload_loaded = true;
  idx->rec = recovery_from_records(records, members);
  idx->have_rec = true;
}
}  // namespace zg

extern "C" int zg_index_load_ex(zg_index *idx, int recover, uint64_t budget, uint32_t memory) {
  try {
    ensure(idx, "null argument", ZG_ERR_INVALID);
    check_query_support(idx->m);
    idx->recover = recover && idx->m.abandoned_recovery;
    if (idx->m.version == 1 && !idx->payload_loaded && !idx->dir.empty()) load_v1(idx);
    if (idx->recover && !idx->have_rec && !idx->dir.empty()) {
      idx->rec = read_recovery_map(idx->dir, idx->m);
      idx->have_rec = true;
    }
    upload_index(idx, budget, memory);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_load(zg_index *idx, int recover, uint64_t budget) {
  return zg_index_load_ex(idx, recover, budget, ZG_MEMORY_ALL);
}

namespace zg {
// The decoded payload of one bank in host memory (decoded from the bank file
// on first use when the index was opened from disk).
const std::vector<uint8_t> &host_payload(zg_index *idx, size_t bank) {
  ensure(bank < idx->m.banks.size(), "bank out of range", ZG_ERR_INVALID);
  if (idx->payload.size() < idx->m.banks.size()) idx->payload.resize(idx->m.banks.size());
  std::vector<uint8_t> &p = idx->payload[bank];
  if (p.size() != idx->m.banks[bank].payload_bytes()) {
    ensure(!idx->dir.empty(), "bank payload not in host memory", ZG_ERR_INVALID);
    if (idx->m.version == 1) {
      std::vector<std::vector<std::pair<uint64_t, uint64_t>>> fallback;
      p = load_v1_bank(idx->dir, idx->m.banks[bank], fallback);
    } else {
      p = load_bank_payload(idx->dir, idx->m.banks[bank], (int)host_threads());
    }
  }
  return p;
}
}  // namespace zg

extern "C" int zg_index_bank_payload(const zg_index *idx, uint32_t bank, const uint8_t **data,
                                     uint64_t *bytes) {
  try {
    ensure(idx && data && bytes, "null argument", ZG_ERR_INVALID);
    const std::vector<uint8_t> &p = host_payload(const_cast<zg_index *>(idx), bank);
    *data = p.data();
    *bytes = p.size();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_manifest(const zg_index *idx, const char **json, uint64_t *len) {
  try {
    ensure(idx && json && len, "null argument", ZG_ERR_INVALID);
    if (idx->manifest_text.empty()) json_write(idx->manifest_text, idx->m.json);
    *json = idx->manifest_text.c_str();
    *len = idx->manifest_text.size();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_recovery(const zg_index *idx, const uint32_t **hashes,
                                 const uint64_t **off, const uint32_t **origins, uint64_t *n) {
  try {
    ensure(idx && hashes && off && origins && n, "null argument", ZG_ERR_INVALID);
    ensure(idx->have_rec, "recovery map not in host memory", ZG_ERR_INVALID);
    *hashes = idx->rec.hashes.data();
    *off = idx->rec.off.data();
    *origins = idx->rec.origins.data();
    *n = idx->rec.size();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_set_tuning(zg_index *idx, const zg_tuning *t) {
  try {
    ensure(idx && t, "null argument", ZG_ERR_INVALID);
    ensure(t->span >= 64 && t->span % 64 == 0, "span must be a positive multiple of 64",
           ZG_ERR_INVALID);
    ensure(t->tiles_per_wave == 1 || t->tiles_per_wave == 2, "tiles_per_wave must be 1 or 2",
           ZG_ERR_INVALID);
    ensure(t->prefetch == 1 || t->prefetch == 2 || t->prefetch == 4 ||
               ((t->prefetch == 8 || t->prefetch == 16) && idx->m.fingerprint_bits <= 4),
           "prefetch must be 1, 2 or 4 (8 or 16 for fingerprints of at most 4 bits)",
           ZG_ERR_INVALID);
    idx->tuning.span = t->span;
    idx->tuning.tiles_per_wave = t->tiles_per_wave;
    idx->tuning.prefetch = t->prefetch;
    idx->tuning.waves_per_cu = t->waves_per_cu;
    idx->tuning.nontemporal = t->nontemporal ? 1 : 0;
    idx->tuning.sweep = t->sweep ? 1 : 0;
    idx->tuning.graph = t->graph ? 1 : 0;
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_index_get_tuning(const zg_index *idx, zg_tuning *t) {
  if (!idx || !t) {
    set_error("null argument");
    return ZG_ERR_INVALID;
  }
  t->span = idx->tuning.span;
  t->tiles_per_wave = idx->tuning.tiles_per_wave;
  t->prefetch = idx->tuning.prefetch;
  t->waves_per_cu = idx->tuning.waves_per_cu;
  t->nontemporal = idx->tuning.nontemporal;
  t->sweep = idx->tuning.sweep;
  t->graph = idx->tuning.graph;
  return 0;
}

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
namespace zg {

static uint32_t max_tiles(const zg_index *idx) {
  uint32_t t = 1;
  for (auto &b : idx->dbanks) t = std::max(t, b.view.n_tiles);
  return t;
}

// Launch shape for the current tuning.
void configure(zg_index *idx, zg_batch *b) {
  const Tuning &t = idx->tuning;
  KernelConfig c{};
  c.bits = (uint32_t)idx->m.fingerprint_bits;
  c.tiles_per_wave = max_tiles(idx) >= 2 ? t.tiles_per_wave : 1;
  c.prefetch = t.prefetch;
  int per_cu = lookup_blocks_per_cu(c);
  if (t.waves_per_cu) per_cu = std::max(1u, std::min((uint32_t)per_cu, (t.waves_per_cu + 3) / 4));
  c.blocks = idx->cu_count * (uint32_t)per_cu;
  b->kcfg = c;
  b->span = t.span;
  b->nontemporal = t.nontemporal;
  // A small batch halves the span until there are at least two work units
  // per wave on the bank with the fewest tile groups, so a short request
  // still spreads over the whole chip.
  uint64_t min_tg = ~0ULL;
  for (const DeviceBank &db : idx->dbanks)
    min_tg = std::min<uint64_t>(min_tg, (db.view.n_tiles + c.tiles_per_wave - 1) / c.tiles_per_wave);
  if (min_tg == ~0ULL) min_tg = 1;
  const uint64_t seeds = b->explicit_keys ? b->n_keys : b->positions;
  const uint64_t waves = (uint64_t)c.blocks * 4;
  while (b->span > 256 && (seeds + b->span - 1) / b->span * min_tg < 2 * waves) b->span /= 2;
}

static void ensure_partials(zg_batch *b) {
  uint64_t waves = (uint64_t)b->kcfg.blocks * 4;
  const uint32_t C = (uint32_t)b->idx->m.colors.size();
  if (!b->records && b->waves < waves) {
    b->partial.alloc(waves * C * sizeof(uint64_t));
    b->waves = waves;
  }
}

LookupArgs lookup_args(zg_index *idx, zg_batch *b, size_t bank) {
  LookupArgs a{};
  a.bank = idx->dbanks[bank].view;
  a.rec = idx->rec_view;
  if (b->flags & ZG_QUERY_NO_RECOVERY) a.rec.enabled = 0;
  a.source = b->explicit_keys ? kSeedKeys : kSeedKmers;
  a.k = (uint32_t)idx->m.k;
  a.codes = b->codes.as<uint64_t>();
  a.invalid = b->invalid.as<uint64_t>();
  a.keys = b->keys.as<uint64_t>();
  a.seg_off = b->explicit_keys ? b->key_off.as<uint64_t>() : b->pos_off.as<uint64_t>();
  a.n_reads = b->n_reads;
  a.begin = 0;
  a.end = b->explicit_keys ? b->n_keys : b->positions;
  uint32_t tiles = a.bank.n_tiles;
  uint32_t T = b->kcfg.tiles_per_wave;
  a.n_tg = (tiles + T - 1) / T;
  a.span = b->span;
  uint64_t spans = (a.end - a.begin + a.span - 1) / a.span;
  a.n_units = spans * a.n_tg;
  ensure(a.n_units < (1ULL << 31), "batch too large for one launch", ZG_ERR_INVALID);
  a.queue = b->queue.as<uint32_t>() + bank;
  a.seed = idx->m.seed;
  a.bits = (uint32_t)idx->m.fingerprint_bits;
  a.records = b->records;
  a.counts = b->counts.as<uint32_t>();
  a.partial = b->partial.as<uint64_t>();
  a.n_colors = (uint32_t)idx->m.colors.size();
  a.nontemporal = b->nontemporal;
  a.sampling = (idx->m.index_mode == 2 && !b->explicit_keys) ? idx->m.modimizer_sampling : 0;
  a.spaced = b->explicit_keys ? SpacedRuns{} : idx->spaced;
  if (idx->m.has_randstrobemers) {
    // both orientations: rows r (forward) and n + r (reverse complement)
    a.records = 1;
    a.n_reads = 2 * b->n_reads;
  }
  if (idx->m.index_mode == 3) {
    // findere: s-mer windows, hits to rows; the scan after the lookups counts
    a.k = (uint32_t)(idx->m.k - idx->m.findere_z);
    a.records = 0;
    a.fd_hits = b->fd_hits.as<uint32_t>();
    a.fd_row_words = idx->fd_row_words;
    a.fd_off = idx->fd_regions[idx->fd_bank_region[bank]].off;
    a.fd_member_bit = idx->fd_bank_member_bit[bank];
  }
  return a;
}

// Zero the outputs of a run (direct launches, before the lookups).
void enqueue_zero_outputs(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  const uint32_t C = (uint32_t)idx->m.colors.size();
  if (idx->m.has_randstrobemers) launch_fill_u32(b->counts.as<uint32_t>(), 2 * b->n_reads * C, 0, s);
  else if (b->records) launch_fill_u32(b->counts.as<uint32_t>(), b->n_reads * C, 0, s);
  else launch_fill_u64(b->partial.as<uint64_t>(), b->waves * C, 0, s);
  launch_fill_u64(b->agg.as<uint64_t>(), C, 0, s);
  if (idx->m.index_mode == 3)
    launch_fill_u32(b->fd_hits.as<uint32_t>(), b->positions * idx->fd_row_words, 0, s);
}

// Findere: the run scan of every hit region into counts / partial sums.
static void enqueue_findere_scan(zg_index *idx, zg_batch *b) {
  for (const zg_index::FdRegion &rg : idx->fd_regions)
    launch_findere_scan(b->fd_hits.as<uint32_t>(), idx->fd_row_words, rg.off, rg.bits,
                        idx->fd_bit_color.as<uint32_t>() + rg.color_at, b->pos_off.as<uint64_t>(),
                        b->n_reads, b->positions, (uint32_t)idx->m.findere_z, b->records ? 1 : 0,
                        b->counts.as<uint32_t>(), b->partial.as<uint64_t>(),
                        (uint32_t)idx->m.colors.size(), b->stream);
}

// The lookups: reset the work counters, one persistent launch per bank. This
// is the part captured in the HIP graph.
// Banks in lookup order: resident ones, then (subindex mode) the streamed one
// already in HBM, then the other streamed ones, each loaded just before its
// lookups (the reference reuses its first loaded bank the same way,
// query_stream.rs:255-265).
static std::vector<size_t> bank_order(const zg_index *idx) {
  std::vector<size_t> order;
  for (size_t i = 0; i < idx->dbanks.size(); i++)
    if (!idx->streamed[i]) order.push_back(i);
  if (idx->stream_holds >= 0) order.push_back((size_t)idx->stream_holds);
  for (size_t i = 0; i < idx->dbanks.size(); i++)
    if (idx->streamed[i] && (int)i != idx->stream_holds) order.push_back(i);
  return order;
}

void enqueue_lookups(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  launch_fill_u32(b->queue.as<uint32_t>(), idx->dbanks.size(), 0, s);
  bool any = b->explicit_keys ? b->n_keys > 0 : b->positions > 0;
  if (any)
    for (size_t i : bank_order(idx)) {
      make_resident(idx, i, s);
      launch_lookup(lookup_args(idx, b, i), b->kcfg, s);
    }
  if (any && idx->m.index_mode == 3) enqueue_findere_scan(idx, b);
  if (idx->m.has_randstrobemers)
    launch_strand_max(b->counts.as<uint32_t>(), b->n_reads * idx->m.colors.size(), s);
}

// Per-colour sums after the lookups.
static void enqueue_reduce(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  const uint32_t C = (uint32_t)idx->m.colors.size();
  if (!b->records && !idx->m.has_randstrobemers)
    launch_reduce_partials(b->partial.as<uint64_t>(), b->waves, C, b->agg.as<uint64_t>(), s);
  else
    launch_colsum(b->counts.as<uint32_t>(), b->n_reads, C, b->agg.as<uint64_t>(), s);
}

static bool sweep_on(const zg_index *idx, const zg_batch *b) {
  (void)b;
  return idx->tuning.sweep && !any_streamed(idx) && sweep_layout_ok(idx);
}

// Frees the sorted sweep's chunk buffers when a batch runs the read-ordered
// lookup (the autotuner may have tried the sweep on it). Findere keeps its
// hit matrix.
static void release_sweep(const zg_index *idx, zg_batch *b) {
  for (DevBuf *d : {&b->sw_h[0], &b->sw_h[1], &b->sw_p[0], &b->sw_p[1], &b->sw_listed,
                    &b->sw_hist, &b->sw_off, &b->sw_start, &b->sw_queue})
    d->reset();
  b->sw_last_p = nullptr;
  b->sw_last_n = nullptr;
  if (idx->m.index_mode != 3) b->fd_hits.reset();
}

static void drop_graph(zg_batch *b) {
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
}

static void drop_sweep_graph(zg_batch *b) {
  if (b->sw_exec) (void)hipGraphExecDestroy(b->sw_exec);
  if (b->sw_graph) (void)hipGraphDestroy(b->sw_graph);
  b->sw_exec = nullptr;
  b->sw_graph = nullptr;
}

// The sweep of a batch as one graph: zeroed outputs, every chunk's sorts,
// staged lookups and sums, the per-colour sums. Captured on the stream after
// an eager run has sized every buffer (nothing allocates during the capture);
// a refused capture leaves the batch on eager launches.
static void capture_sweep(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  bool ok = false;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    bool enqueued = true;
    b->sw_capturing = true;
    try {
      enqueue_zero_outputs(idx, b);
      sweep_run(idx, b);
      enqueue_reduce(idx, b);
    } catch (...) {
      enqueued = false;
    }
    b->sw_capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &g);
    if (enqueued && end == hipSuccess && g) {
      b->sw_graph = g;
      ok = hipGraphInstantiate(&b->sw_exec, g, nullptr, nullptr, 0) == hipSuccess;
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();
  }
  if (!ok) {
    drop_sweep_graph(b);
    b->sw_graph_failed = true;
  }
}

// Milliseconds between two completed events; -1 when the runtime cannot time
// them (timing never fails a query).
double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return -1.0;
  }
  return ms;
}

// Randstrobemer seeds (anchors::scan, src/anchors.rs:284-351) for both query
// orientations (record_query.rs:659-685): explicit keys, forward ones first.
static void seed_strobes(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  const Manifest &m = idx->m;
  const uint64_t words = (b->positions + 63) / 64;
  const uint32_t len = (uint32_t)(m.index_mode == 1 ? m.minimizer_size : m.k);
  b->rs_cand.reserve((words + 1) * 8);
  b->rs_counts.reserve((words + 1) * 4);
  b->rs_off.reserve((words + 1) * 8);
  b->rs_inv_off.reserve((words + 1) * 8);
  const size_t tb = scan_temp_bytes_u32_to_u64(words + 1);
  b->scan_temp.reserve(tb);
  // run ids: invalid positions before each word
  launch_popc_words(b->invalid.as<uint64_t>(), words, b->rs_counts.as<uint32_t>(), s);
  launch_fill_u32(b->rs_counts.as<uint32_t>() + words, 1, 0, s);
  launch_exclusive_scan_u32_to_u64(b->rs_counts.as<uint32_t>(), b->rs_inv_off.as<uint64_t>(),
                                   words + 1, b->scan_temp.p, tb, s);
  if (m.index_mode == 1) {
    const uint64_t blocks = (b->positions + ocm_block_positions() - 1) / ocm_block_positions();
    b->flag.reserve(b->positions + 64);
    b->delta.reserve(b->positions + 64);
    b->block_counts.reserve((blocks + 1) * sizeof(uint32_t));
    launch_ocm_select(b->codes.as<uint64_t>(), b->invalid.as<uint64_t>(), b->pos_off.as<uint64_t>(),
                      b->n_reads, b->positions, (uint32_t)m.k, (uint32_t)m.minimizer_size, m.seed,
                      idx->inner_table.as<uint64_t>(), b->flag.as<uint8_t>(),
                      b->delta.as<uint8_t>(), b->block_counts.as<uint32_t>(), nullptr, s);
    launch_fill_u64(b->rs_cand.as<uint64_t>(), words + 1, 0, s);
    launch_strobe_mark_selected(b->flag.as<uint8_t>(), b->delta.as<uint8_t>(), b->positions,
                                b->rs_cand.as<uint64_t>(), s);
    launch_popc_words(b->rs_cand.as<uint64_t>(), words, b->rs_counts.as<uint32_t>(), s);
  } else {
    launch_strobe_kmer_candidates(b->invalid.as<uint64_t>(), b->positions, len,
                                  b->rs_cand.as<uint64_t>(), b->rs_counts.as<uint32_t>(), s);
  }
  launch_fill_u32(b->rs_counts.as<uint32_t>() + words, 1, 0, s);
  launch_exclusive_scan_u32_to_u64(b->rs_counts.as<uint32_t>(), b->rs_off.as<uint64_t>(), words + 1,
                                   b->scan_temp.p, tb, s);
  uint64_t total = 0;
  check_hip(hipMemcpyAsync(&total, b->rs_off.as<uint64_t>() + words, 8, hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "seed");
  b->key_win.reserve((total + 64) * 8);
  b->rs_hash.reserve((total + 64) * 8);
  b->rs_run.reserve((total + 64) * 8);
  b->keys.reserve((2 * total + 64) * 8);
  launch_strobe_compact(b->rs_cand.as<uint64_t>(), b->rs_off.as<uint64_t>(), b->invalid.as<uint64_t>(),
                        b->rs_inv_off.as<uint64_t>(), b->codes.as<uint64_t>(), b->positions, len,
                        idx->spaced, b->key_win.as<uint64_t>(), b->rs_hash.as<uint64_t>(),
                        b->rs_run.as<uint64_t>(), s);
  launch_strobe_link(b->key_win.as<uint64_t>(), b->rs_hash.as<uint64_t>(), b->rs_run.as<uint64_t>(),
                     total, m.strobe_min, m.strobe_max, b->keys.as<uint64_t>(), s);
  b->key_off.reserve((2 * b->n_reads + 1) * 8);
  launch_segment_bounds(b->key_win.as<uint64_t>(), total, b->pos_off.as<uint64_t>(), b->n_reads,
                        b->key_off.as<uint64_t>(), s);
  launch_strobe_offsets(b->key_off.as<uint64_t>(), b->n_reads, total, s);
  launch_diff(b->key_off.as<uint64_t>(), b->n_reads, b->totals.as<uint64_t>(), s);
  check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
  b->h_seg_off.assign(2 * b->n_reads + 1, 0);
  check_hip(hipMemcpyAsync(b->h_seg_off.data(), b->key_off.p, (2 * b->n_reads + 1) * 8,
                           hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "seed");
  b->n_keys = 2 * total;
  b->rs_forward = total;
  b->explicit_keys = true;
}

// Long k-mers (k > 32) as explicit keys: hashed per window, then compacted
// in position order (modimizers keep only the sampled ones).
static void seed_long_kmers(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  const Manifest &m = idx->m;
  const uint64_t words = (b->positions + 63) / 64;
  b->rs_cand.reserve((words + 1) * 8);
  b->rs_counts.reserve((words + 1) * 4);
  b->rs_off.reserve((words + 1) * 8);
  b->rs_hash.reserve((b->positions + 64) * 8);
  const size_t tb = scan_temp_bytes_u32_to_u64(words + 1);
  b->scan_temp.reserve(tb);
  launch_long_kmer_keys(b->codes.as<uint64_t>(), b->invalid.as<uint64_t>(), b->positions,
                        long_kmer_params((uint32_t)m.k), m.seed,
                        m.index_mode == 2 ? m.modimizer_sampling : 0, b->rs_cand.as<uint64_t>(),
                        b->rs_counts.as<uint32_t>(), b->rs_hash.as<uint64_t>(), s);
  launch_fill_u32(b->rs_counts.as<uint32_t>() + words, 1, 0, s);
  launch_exclusive_scan_u32_to_u64(b->rs_counts.as<uint32_t>(), b->rs_off.as<uint64_t>(), words + 1,
                                   b->scan_temp.p, tb, s);
  uint64_t total = 0;
  check_hip(hipMemcpyAsync(&total, b->rs_off.as<uint64_t>() + words, 8, hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "seed");
  b->keys.reserve((total + 64) * 8);
  b->key_win.reserve((total + 64) * 8);
  launch_compact_kept(b->rs_cand.as<uint64_t>(), b->rs_off.as<uint64_t>(), b->rs_hash.as<uint64_t>(),
                      b->positions, b->keys.as<uint64_t>(), b->key_win.as<uint64_t>(), s);
  b->key_off.reserve((b->n_reads + 1) * 8);
  launch_segment_bounds(b->key_win.as<uint64_t>(), total, b->pos_off.as<uint64_t>(), b->n_reads,
                        b->key_off.as<uint64_t>(), s);
  launch_diff(b->key_off.as<uint64_t>(), b->n_reads, b->totals.as<uint64_t>(), s);
  check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
  b->h_seg_off.assign(b->n_reads + 1, 0);
  check_hip(hipMemcpyAsync(b->h_seg_off.data(), b->key_off.p, (b->n_reads + 1) * 8,
                           hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "seed");
  b->n_keys = total;
  b->explicit_keys = true;
  b->keys_at_position = true;
}

void seed_batch(zg_index *idx, zg_batch *b) {
  hipStream_t s = b->stream;
  const Manifest &m = idx->m;
  check_hip(hipEventRecord(b->ev[0], s), "hipEventRecord");
  const uint64_t cwords = code_words(b->positions, m.k), iwords = invalid_words(b->positions, m.k);
  launch_fill_u64(b->codes.as<uint64_t>(), cwords, 0, s);
  launch_fill_u64(b->invalid.as<uint64_t>(), iwords, ~0ULL, s);
  if (m.homopolymer_compression) {
    // each read compressed in place: its bases first, the rest invalid
    const uint64_t words = (b->positions + 63) / 64;
    b->hpc_keep.reserve((words + 1) * 8);
    b->hpc_counts.reserve((words + 1) * 4);
    b->hpc_off.reserve((words + 1) * 8);
    launch_hpc_keep(b->seq.as<uint8_t>(), b->read_off.as<uint64_t>(), b->pos_off.as<uint64_t>(),
                    b->n_reads, b->positions, b->hpc_keep.as<uint64_t>(),
                    b->hpc_counts.as<uint32_t>(), s);
    launch_fill_u32(b->hpc_counts.as<uint32_t>() + words, 1, 0, s);
    const size_t tb = scan_temp_bytes_u32_to_u64(words + 1);
    b->scan_temp.reserve(tb);
    launch_exclusive_scan_u32_to_u64(b->hpc_counts.as<uint32_t>(), b->hpc_off.as<uint64_t>(),
                                     words + 1, b->scan_temp.p, tb, s);
    launch_hpc_place(b->seq.as<uint8_t>(), b->read_off.as<uint64_t>(), b->pos_off.as<uint64_t>(),
                     b->n_reads, b->positions, b->hpc_keep.as<uint64_t>(),
                     b->hpc_off.as<uint64_t>(), b->codes.as<uint64_t>(),
                     b->invalid.as<uint64_t>(), s);
  } else {
    launch_pack(b->seq.as<uint8_t>(), b->read_off.as<uint64_t>(), b->pos_off.as<uint64_t>(),
                b->n_reads, b->positions, b->codes.as<uint64_t>(), b->invalid.as<uint64_t>(), s);
  }
  launch_fill_u64(b->totals.as<uint64_t>(), b->n_reads, 0, s);
  b->keys_at_position = false;
  if (m.has_randstrobemers) {
    seed_strobes(idx, b);
    b->last_seed_ms = elapsed_ms(b->ev[0], b->ev[1]);
    return;
  }
  if ((m.index_mode == 0 || m.index_mode == 2) && m.k > 32) {
    seed_long_kmers(idx, b);
    b->last_seed_ms = elapsed_ms(b->ev[0], b->ev[1]);
    return;
  }
  if (m.index_mode == 3) b->fd_hits.reserve(std::max<uint64_t>(1, b->positions * idx->fd_row_words) * 4);
  if (m.index_mode != 1) {
    // k-mers, modimizers (sampled k-mers), or findere (its total is the
    // number of valid k-mer windows, Scorer::sequence)
    launch_count_kmers(b->invalid.as<uint64_t>(), b->codes.as<uint64_t>(), m.seed,
                       m.index_mode == 2 ? m.modimizer_sampling : 0, b->pos_off.as<uint64_t>(),
                       b->n_reads, b->positions, (uint32_t)m.k, b->totals.as<uint64_t>(), s);
    check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
    check_hip(hipEventSynchronize(b->ev[1]), "seed");
  } else {
    const uint64_t blocks = (b->positions + ocm_block_positions() - 1) / ocm_block_positions();
    b->flag.reserve(b->positions + 64);
    b->delta.reserve(b->positions + 64);
    b->block_counts.reserve((blocks + 1) * sizeof(uint32_t));
    b->block_off.reserve((blocks + 1) * sizeof(uint64_t));
    launch_ocm_select(b->codes.as<uint64_t>(), b->invalid.as<uint64_t>(), b->pos_off.as<uint64_t>(),
                      b->n_reads, b->positions, (uint32_t)m.k, (uint32_t)m.minimizer_size, m.seed,
                      idx->inner_table.as<uint64_t>(), b->flag.as<uint8_t>(),
                      b->delta.as<uint8_t>(), b->block_counts.as<uint32_t>(), nullptr, s);
    launch_fill_u32(b->block_counts.as<uint32_t>() + blocks, 1, 0, s);
    size_t tb = scan_temp_bytes_u32_to_u64(blocks + 1);
    b->scan_temp.reserve(tb);
    launch_exclusive_scan_u32_to_u64(b->block_counts.as<uint32_t>(), b->block_off.as<uint64_t>(),
                                     blocks + 1, b->scan_temp.p, tb, s);
    uint64_t total = 0;
    check_hip(hipMemcpyAsync(&total, b->block_off.as<uint64_t>() + blocks, 8, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "seed");
    b->n_keys = total;
    b->keys.reserve((total + 64) * 8);
    b->key_win.reserve((total + 64) * 8);
    launch_ocm_compact(b->codes.as<uint64_t>(), b->flag.as<uint8_t>(), b->delta.as<uint8_t>(),
                       b->block_off.as<uint64_t>(), b->positions, (uint32_t)m.minimizer_size,
                       idx->spaced, b->keys.as<uint64_t>(), b->key_win.as<uint64_t>(), s);
    b->key_off.reserve((b->n_reads + 1) * 8);
    launch_segment_bounds(b->key_win.as<uint64_t>(), total, b->pos_off.as<uint64_t>(), b->n_reads,
                          b->key_off.as<uint64_t>(), s);
    launch_diff(b->key_off.as<uint64_t>(), b->n_reads, b->totals.as<uint64_t>(), s);
    check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
    b->h_seg_off.assign(b->n_reads + 1, 0);
    check_hip(hipMemcpyAsync(b->h_seg_off.data(), b->key_off.p, (b->n_reads + 1) * 8,
                             hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "seed");
    b->explicit_keys = true;
  }
  b->last_seed_ms = elapsed_ms(b->ev[0], b->ev[1]);
}

}  // namespace zg

extern "C" int zg_batch_create(zg_index *idx, const uint8_t *seq, const uint64_t *read_off,
                               uint64_t n_reads, uint32_t flags, zg_batch **out) {
  try {
    ensure(idx && out && (n_reads == 0 || read_off), "null argument", ZG_ERR_INVALID);
    ensure(idx->on_device, "index is not resident on the device (call zg_index_load)",
           ZG_ERR_INVALID);
    auto b = std::make_unique<zg_batch>();
    b->idx = idx;
    b->flags = flags;
    b->records = (flags & ZG_QUERY_RECORDS) != 0;
    b->n_reads = n_reads;
    b->stream = idx->stream;
    hipStream_t s = b->stream;
    for (auto &e : b->ev) check_hip(hipEventCreate(&e), "hipEventCreate");
    uint64_t r0 = n_reads ? read_off[0] : 0;
    b->bases = n_reads ? read_off[n_reads] - r0 : 0;
    ensure(b->bases == 0 || seq, "null sequence", ZG_ERR_INVALID);
    b->positions = b->bases + n_reads;
    std::vector<uint64_t> roff(n_reads + 1), poff(n_reads + 1);
    for (uint64_t r = 0; r <= n_reads; r++) {
      roff[r] = read_off[r] - r0;
      ensure(r == 0 || roff[r] >= roff[r - 1], "read offsets must be non-decreasing",
             ZG_ERR_INVALID);
      poff[r] = roff[r] + r;
    }
    b->seq.alloc(b->bases + 64);
    if (b->bases)
      check_hip(hipMemcpyAsync(b->seq.p, seq + r0, b->bases, hipMemcpyHostToDevice, s),
                "hipMemcpyAsync");
    upload(b->read_off, roff, s);
    upload(b->pos_off, poff, s);
    b->codes.alloc(code_words(b->positions, idx->m.k) * 8);
    b->invalid.alloc(invalid_words(b->positions, idx->m.k) * 8);
    b->totals.alloc((n_reads + 1) * 8);
    const uint32_t C = (uint32_t)idx->m.colors.size();
    if (idx->m.has_randstrobemers) b->counts.alloc(std::max<uint64_t>(1, 2 * n_reads * C) * 4);
    else if (b->records) b->counts.alloc(std::max<uint64_t>(1, n_reads * C) * 4);
    b->agg.alloc((uint64_t)C * 8 + 8);
    b->queue.alloc((idx->dbanks.size() + 1) * 4);
    b->h_seg_off = poff;
    zg::seed_batch(idx, b.get());
    zg::configure(idx, b.get());
    zg::ensure_partials(b.get());
    *out = b.release();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_reseed(zg_index *idx, zg_batch *b) {
  try {
    ensure(idx && b && b->idx == idx, "invalid batch", ZG_ERR_INVALID);
    zg::seed_batch(idx, b);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_run(zg_index *idx, zg_batch *b) {
  try {
    ensure(idx && b && b->idx == idx, "invalid batch", ZG_ERR_INVALID);
    hipStream_t s = b->stream;
    if (zg::sweep_on(idx, b)) {
      zg::drop_graph(b);
      if (!idx->tuning.graph) zg::drop_sweep_graph(b);
      if (b->sw_exec) {
        // one launch; only the whole run is timed
        check_hip(hipEventRecord(b->ev[0], s), "hipEventRecord");
        check_hip(hipGraphLaunch(b->sw_exec, s), "hipGraphLaunch");
        check_hip(hipEventRecord(b->ev[3], s), "hipEventRecord");
        check_hip(hipStreamSynchronize(s), "lookup");
        b->last_total_ms = b->last_lookup_ms = zg::elapsed_ms(b->ev[0], b->ev[3]);
        b->sweep_sort_ms = b->sweep_lookup_ms = b->sweep_sum_ms = b->sweep_recovery_ms = -1;
        b->sweep_folded = zg::sweep_folded(b);
        return 0;
      }
      check_hip(hipEventRecord(b->ev[0], s), "hipEventRecord");
      zg::enqueue_zero_outputs(idx, b);
      check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
      zg::sweep_run(idx, b);
      check_hip(hipEventRecord(b->ev[2], s), "hipEventRecord");
      zg::enqueue_reduce(idx, b);
      check_hip(hipEventRecord(b->ev[3], s), "hipEventRecord");
      check_hip(hipStreamSynchronize(s), "lookup");
      b->last_lookup_ms = zg::elapsed_ms(b->ev[1], b->ev[2]);
      b->last_total_ms = zg::elapsed_ms(b->ev[0], b->ev[3]);
      zg::sweep_times(b, b->sweep_sort_ms, b->sweep_lookup_ms, b->sweep_recovery_ms,
                      b->sweep_sum_ms);
      b->sweep_folded = zg::sweep_folded(b);
      if (idx->tuning.graph && !b->sw_graph_failed) zg::capture_sweep(idx, b);
      return 0;
    }
    zg::release_sweep(idx, b);
    // banks streamed per query cannot sit in a graph: launch directly
    if (any_streamed(idx)) zg::drop_graph(b);
    if (!b->graph_exec && !b->graph_failed && !any_streamed(idx)) {
      // Capture the lookups once; replays reuse the instantiated graph. If any
      // step of the capture is refused, fall back to direct launches.
      bool captured = false;
      if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        bool enqueued = true;
        try {
          zg::enqueue_lookups(idx, b);
        } catch (...) {
          enqueued = false;
        }
        hipGraph_t g = nullptr;
        hipError_t end = hipStreamEndCapture(s, &g);
        if (enqueued && end == hipSuccess && g) {
          b->graph = g;
          if (hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0) == hipSuccess)
            captured = true;
        } else if (g) {
          (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
      }
      if (!captured) {
        zg::drop_graph(b);
        b->graph_failed = true;
      }
    }
    check_hip(hipEventRecord(b->ev[0], s), "hipEventRecord");
    zg::enqueue_zero_outputs(idx, b);
    check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
    if (b->graph_exec) check_hip(hipGraphLaunch(b->graph_exec, s), "hipGraphLaunch");
    else zg::enqueue_lookups(idx, b);
    check_hip(hipEventRecord(b->ev[2], s), "hipEventRecord");
    zg::enqueue_reduce(idx, b);
    check_hip(hipEventRecord(b->ev[3], s), "hipEventRecord");
    check_hip(hipStreamSynchronize(s), "lookup");
    b->last_lookup_ms = zg::elapsed_ms(b->ev[1], b->ev[2]);
    b->last_total_ms = zg::elapsed_ms(b->ev[0], b->ev[3]);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_run_plain(zg_index *idx, zg_batch *b, uint64_t first, uint64_t count) {
  try {
    ensure(idx && b && b->idx == idx, "invalid batch", ZG_ERR_INVALID);
    ensure(first <= b->n_reads && count <= b->n_reads - first, "read range out of bounds",
           ZG_ERR_INVALID);
    hipStream_t s = b->stream;
    const uint32_t C = (uint32_t)idx->m.colors.size();
    const size_t nb = idx->dbanks.size();
    zg::DevBuf queues;
    const bool strobes = idx->m.has_randstrobemers;
    const uint64_t strands = strobes ? 2 : 1;
    queues.alloc((strands * count * nb + 1) * 4);
    check_hip(hipEventRecord(b->ev[0], s), "hipEventRecord");
    if (strobes) zg::launch_fill_u32(b->counts.as<uint32_t>(), 2 * b->n_reads * C, 0, s);
    else if (b->records) zg::launch_fill_u32(b->counts.as<uint32_t>(), b->n_reads * C, 0, s);
    else zg::launch_fill_u64(b->partial.as<uint64_t>(), b->waves * C, 0, s);
    zg::launch_fill_u64(b->agg.as<uint64_t>(), C, 0, s);
    zg::launch_fill_u32(queues.as<uint32_t>(), strands * count * nb, 0, s);
    if (idx->m.index_mode == 3)
      zg::launch_fill_u32(b->fd_hits.as<uint32_t>(), b->positions * idx->fd_row_words, 0, s);
    check_hip(hipEventRecord(b->ev[1], s), "hipEventRecord");
    for (size_t i : zg::bank_order(idx)) {
      zg::make_resident(idx, i, s);
      for (uint64_t q = 0; q < strands * count; q++) {
        // segment of read first + q (or, for randstrobes, its reverse complement)
        const uint64_t r = q < count ? first + q : b->n_reads + first + (q - count);
        uint64_t lo = b->h_seg_off[r], hi = b->h_seg_off[r + 1];
        if (hi <= lo) continue;
        zg::LookupArgs a = zg::lookup_args(idx, b, i);
        a.begin = lo;
        a.end = hi;
        a.span = 64;  // one read per launch: the smallest unit spreads it widest
        uint64_t spans = (hi - lo + a.span - 1) / a.span;
        a.n_units = spans * a.n_tg;
        a.queue = queues.as<uint32_t>() + q * nb + i;
        zg::KernelConfig c = b->kcfg;
        uint64_t want = (a.n_units + 3) / 4;
        c.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, b->kcfg.blocks));
        zg::launch_lookup(a, c, s);
      }
    }
    if (idx->m.index_mode == 3) zg::enqueue_findere_scan(idx, b);
    if (strobes) zg::launch_strand_max(b->counts.as<uint32_t>(), b->n_reads * C, s);
    check_hip(hipEventRecord(b->ev[2], s), "hipEventRecord");
    if (!b->records && !strobes)
      zg::launch_reduce_partials(b->partial.as<uint64_t>(), b->waves, C, b->agg.as<uint64_t>(), s);
    else
      zg::launch_colsum(b->counts.as<uint32_t>(), b->n_reads, C, b->agg.as<uint64_t>(), s);
    check_hip(hipEventRecord(b->ev[3], s), "hipEventRecord");
    check_hip(hipStreamSynchronize(s), "plain lookup");
    b->last_lookup_ms = zg::elapsed_ms(b->ev[1], b->ev[2]);
    b->last_total_ms = zg::elapsed_ms(b->ev[0], b->ev[3]);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_results(zg_batch *b, uint64_t *totals, uint32_t *counts,
                                uint64_t *aggregate) {
  try {
    ensure(b, "null argument", ZG_ERR_INVALID);
    hipStream_t s = b->stream;
    const uint32_t C = (uint32_t)b->idx->m.colors.size();
    if (totals && b->n_reads)
      check_hip(hipMemcpyAsync(totals, b->totals.p, b->n_reads * 8, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    if (counts) {
      ensure(b->records, "per-read counts need ZG_QUERY_RECORDS", ZG_ERR_INVALID);
      if (b->n_reads && C)
        check_hip(hipMemcpyAsync(counts, b->counts.p, b->n_reads * C * 4ULL, hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync");
    }
    if (aggregate && C)
      check_hip(hipMemcpyAsync(aggregate, b->agg.p, C * 8ULL, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "results");
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_get_stats(const zg_batch *b, zg_batch_stats *o) {
  if (!b || !o) {
    set_error("null argument");
    return ZG_ERR_INVALID;
  }
  try {
    memset(o, 0, sizeof *o);
    const zg_index *idx = b->idx;
    o->n_reads = b->n_reads;
    o->bases = b->bases;
    std::vector<uint64_t> totals(b->n_reads);
    if (b->n_reads)
      check_hip(hipMemcpy(totals.data(), b->totals.p, b->n_reads * 8, hipMemcpyDeviceToHost),
                "hipMemcpy");
    for (uint64_t t : totals) o->seeds += t;
    uint64_t per_lookup = 0;
    for (auto &bk : idx->m.banks) {
      uint64_t G = bk.groups.size();
      uint64_t parts = bk.routed ? bk.partitions : 1;
      // every seed probes one partition of each class
      per_lookup += idx->m.hashes * ((G * idx->m.fingerprint_bits + 7) / 8) / parts;
    }
    o->lookups = o->seeds * idx->m.classes();
    o->algorithmic_bytes = o->seeds * per_lookup;
    o->last_lookup_ms = b->last_lookup_ms;
    o->last_seed_ms = b->last_seed_ms;
    o->last_total_ms = b->last_total_ms;
    o->tuning.span = b->span;
    o->tuning.tiles_per_wave = b->kcfg.tiles_per_wave;
    o->tuning.prefetch = b->kcfg.prefetch;
    o->tuning.waves_per_cu = b->kcfg.blocks * 4 / std::max(1u, idx->cu_count);
    o->tuning.nontemporal = b->nontemporal;
    o->tuning.sweep = zg::sweep_on(idx, b) ? 1 : 0;
    o->tuning.graph = o->tuning.sweep && b->sw_exec ? 1 : 0;
    o->sweep_folded = o->tuning.sweep ? b->sweep_folded : 0;
    o->sweep_launches = o->tuning.sweep ? b->sweep_chunks : 0;
    o->sweep_lookup_ms = o->tuning.sweep ? b->sweep_lookup_ms : 0;
    o->sweep_sort_ms = o->tuning.sweep ? b->sweep_sort_ms : 0;
    o->sweep_sum_ms = o->tuning.sweep ? b->sweep_sum_ms : 0;
    o->sweep_recovery_ms = o->tuning.sweep ? b->sweep_recovery_ms : 0;
    if (o->tuning.sweep) {
      uint64_t bank_bytes = 0;
      for (auto &bk : idx->m.banks) bank_bytes += device_bytes(bk);
      o->sweep_stage_launches = b->sweep_chunks * idx->m.banks.size();
      o->sweep_hit_row_bytes = idx->sw_row_words * 4ULL;
      o->sweep_algorithmic_bytes =
          b->sweep_chunks * bank_bytes + o->seeds * (12 + o->sweep_hit_row_bytes);
      o->sweep_plane_banks = 0;
      for (const SweepClass &c : idx->sw_classes)
        if (c.planes) o->sweep_plane_banks += c.banks.size();
    }
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_store_probe(zg_batch *b, double *gbs) {
  try {
    ensure(b && gbs, "null argument", ZG_ERR_INVALID);
    *gbs = zg::sweep_store_probe(b);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_batch_destroy(zg_batch *b) {
  delete b;
  return 0;
}

extern "C" int zg_autotune(zg_index *idx, zg_batch *b, zg_tuning *chosen) {
  try {
    ensure(idx && b && b->idx == idx, "invalid batch", ZG_ERR_INVALID);
    zg::Tuning best = idx->tuning;
    double best_ms = 1e30;
    std::vector<uint32_t> tiles = {1};
    if (zg::max_tiles(idx) >= 2) tiles.push_back(2);
    int base_per_cu = 0;
    const bool wide = idx->m.fingerprint_bits >= 8;  // 16-byte lane loads
    std::vector<uint32_t> depths = {1, 2, 4};
    if (idx->m.fingerprint_bits <= 4) depths.insert(depths.end(), {8, 16});
    // the sorted sweep, where the index and batch allow it, with shorter spans
    // (neighbouring waves then share rows)
    const bool can_sweep = !any_streamed(idx) && zg::sweep_layout_ok(idx);
    for (uint32_t sw : {0u, 1u})
    for (uint32_t span : {1024u, 4096u, 16384u})
      for (uint32_t t : tiles)
        for (uint32_t d : depths)
          for (uint32_t frac : {1u, 2u})
            for (uint32_t nt : {0u, 1u}) {
            if (nt && !wide) continue;
            // the sweep's staged lookup has one launch shape of its own
            if (sw && (!can_sweep || span != 1024u || t != 1u || d != 1u || frac != 1u || nt))
              continue;
            zg::Tuning cand;
            cand.sweep = sw;
            cand.span = span;
            cand.tiles_per_wave = t;
            cand.prefetch = d;
            cand.nontemporal = nt;
            zg::KernelConfig kc{(uint32_t)idx->m.fingerprint_bits, t, d, 0};
            base_per_cu = zg::lookup_blocks_per_cu(kc);
            cand.waves_per_cu = (uint32_t)std::max(1, base_per_cu / (int)frac) * 4;
            if (frac == 2 && base_per_cu < 2) continue;
            idx->tuning = cand;
            zg::drop_graph(b);
            zg::configure(idx, b);
            zg::ensure_partials(b);
            double ms = 1e30;
            for (int rep = 0; rep < 2; rep++) {
              int rc = zg_batch_run(idx, b);
              if (rc && sw) {  // the sweep cannot run here (device memory): not a candidate
                ms = 1e30;
                break;
              }
              if (rc) return rc;
              ms = std::min(ms, b->last_lookup_ms);
            }
            if (getenv("ZG_AUTOTUNE_TRACE"))
              fprintf(stderr,
                      "autotune sweep=%u span=%u tiles=%u prefetch=%u waves/cu=%u nt=%u: %.3f ms\n",
                      cand.sweep, cand.span, cand.tiles_per_wave, cand.prefetch,
                      cand.waves_per_cu, cand.nontemporal, ms);
            if (ms < best_ms) {
              best_ms = ms;
              best = cand;
            }
          }
    idx->tuning = best;
    zg::drop_graph(b);
    zg::configure(idx, b);
    zg::ensure_partials(b);
    if (chosen) zg_index_get_tuning(idx, chosen);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_gather_probe(zg_index *idx, uint32_t bank, uint32_t row_bytes,
                               uint64_t n_rows, double *gbs) {
  try {
    ensure(idx && gbs, "null argument", ZG_ERR_INVALID);
    ensure(idx->on_device, "index is not resident on the device (call zg_index_load)",
           ZG_ERR_INVALID);
    ensure(bank < idx->dbanks.size(), "bank out of range", ZG_ERR_INVALID);
    make_resident(idx, bank, idx->stream);
    const DeviceBank &db = idx->dbanks[bank];
    const uint64_t bytes = device_bytes(idx->m.banks[bank]);
    ensure(n_rows > 0 && bytes >= row_bytes, "empty probe", ZG_ERR_INVALID);
    hipStream_t s = idx->stream;
    DevBuf sink;
    sink.alloc(16);
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "hipEventCreate");
    check_hip(hipEventCreate(&e1), "hipEventCreate");
    const uint32_t blocks = idx->cu_count * 8;  // 32 waves per CU
    const uint8_t *data = db.view.data;
    bool ok = launch_gather_probe(data, bytes, row_bytes, n_rows / 8, blocks, sink.as<uint32_t>(),
                                  s);  // warm-up
    if (ok) {
      check_hip(hipEventRecord(e0, s), "hipEventRecord");
      launch_gather_probe(data, bytes, row_bytes, n_rows, blocks, sink.as<uint32_t>(), s);
      check_hip(hipEventRecord(e1, s), "hipEventRecord");
    }
    check_hip(hipStreamSynchronize(s), "gather probe");
    double ms = ok ? elapsed_ms(e0, e1) : -1;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ensure(ok, "row_bytes must be 512, 1024, 2048, 4096 or 8192", ZG_ERR_INVALID);
    *gbs = ms > 0 ? (double)n_rows * row_bytes / (ms * 1e6) : 0.0;
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_gather_probe_keys(zg_index *idx, uint32_t bank, uint64_t n_keys, double *gbs) {
  try {
    ensure(idx && gbs, "null argument", ZG_ERR_INVALID);
    ensure(idx->on_device, "index is not resident on the device (call zg_index_load)",
           ZG_ERR_INVALID);
    ensure(bank < idx->dbanks.size(), "bank out of range", ZG_ERR_INVALID);
    ensure(n_keys > 0, "empty probe", ZG_ERR_INVALID);
    make_resident(idx, bank, idx->stream);
    const BankDev &v = idx->dbanks[bank].view;
    hipStream_t s = idx->stream;
    DevBuf sink;
    sink.alloc(16);
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "hipEventCreate");
    check_hip(hipEventCreate(&e1), "hipEventCreate");
    const uint32_t blocks = idx->cu_count * 8;
    bool ok = launch_gather_keys(v.data, v.row_bytes, v.geo, v.hashes, n_keys / 8, blocks,
                                 sink.as<uint32_t>(), s);  // warm-up
    if (ok) {
      check_hip(hipEventRecord(e0, s), "hipEventRecord");
      launch_gather_keys(v.data, v.row_bytes, v.geo, v.hashes, n_keys, blocks, sink.as<uint32_t>(), s);
      check_hip(hipEventRecord(e1, s), "hipEventRecord");
    }
    check_hip(hipStreamSynchronize(s), "gather probe");
    double ms = ok ? elapsed_ms(e0, e1) : -1;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ensure(ok, "key probe needs rows of 512, 1024 or 2048 bytes", ZG_ERR_INVALID);
    *gbs = ms > 0 ? (double)n_keys * v.hashes * v.row_bytes / (ms * 1e6) : 0.0;
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

// ---------------------------------------------------------------------------
// Host-buffer queries
// ---------------------------------------------------------------------------
extern "C" int zg_query_reads(zg_index *idx, const uint8_t *seq, const uint64_t *read_off,
                              uint64_t n_reads, uint32_t flags, uint64_t *totals,
                              uint32_t *counts, uint64_t *aggregate) {
  try {
    ensure(idx && (n_reads == 0 || read_off), "null argument", ZG_ERR_INVALID);
    ensure(!counts || (flags & ZG_QUERY_RECORDS), "per-read counts need ZG_QUERY_RECORDS",
           ZG_ERR_INVALID);
    const uint32_t C = (uint32_t)idx->m.colors.size();
    if (aggregate) memset(aggregate, 0, C * 8ULL);
    std::vector<uint64_t> agg(C);
    const uint64_t max_bases = 1ULL << 30;
    const uint64_t max_cells = 1ULL << 30;  // per-read counter cells per chunk
    uint64_t r = 0;
    while (r < n_reads || (n_reads == 0 && r == 0)) {
      uint64_t e = r;
      while (e < n_reads && (e == r || (read_off[e + 1] - read_off[r] <= max_bases &&
                                        (!counts || (e + 1 - r) * (uint64_t)C <= max_cells))))
        e++;
      zg_batch *b = nullptr;
      int rc = zg_batch_create(idx, seq, read_off + r, e - r, flags, &b);
      if (rc) return rc;
      rc = zg_batch_run(idx, b);
      if (!rc)
        rc = zg_batch_results(b, totals ? totals + r : nullptr,
                              counts ? counts + r * (uint64_t)C : nullptr,
                              aggregate ? agg.data() : nullptr);
      zg_batch_destroy(b);
      if (rc) return rc;
      if (aggregate)
        for (uint32_t c = 0; c < C; c++) aggregate[c] += agg[c];
      if (n_reads == 0) break;
      r = e;
    }
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_extract_seeds(zg_index *idx, const uint8_t *seq, const uint64_t *read_off,
                                uint64_t n_reads, uint64_t *key_off, uint64_t *keys,
                                uint64_t *positions, uint64_t cap) {
  try {
    ensure(idx && key_off && (n_reads == 0 || read_off), "null argument", ZG_ERR_INVALID);
    zg_batch *b = nullptr;
    int rc = zg_batch_create(idx, seq, read_off, n_reads, 0, &b);
    if (rc) return rc;
    std::unique_ptr<zg_batch> hold(b);
    hipStream_t s = idx->stream;
    if (idx->m.has_randstrobemers) {
      // the forward keys, one per anchor, at the anchor's start
      const uint64_t N = b->rs_forward;
      std::vector<uint64_t> start(N), kk(N);
      if (N) {
        check_hip(hipMemcpyAsync(start.data(), b->key_win.p, N * 8, hipMemcpyDeviceToHost, s), "copy");
        check_hip(hipMemcpyAsync(kk.data(), b->keys.p, N * 8, hipMemcpyDeviceToHost, s), "copy");
      }
      check_hip(hipStreamSynchronize(s), "copy");
      for (uint64_t r = 0; r <= n_reads; r++) key_off[r] = b->h_seg_off[r];
      if (N > cap) fail(ZG_ERR_INVALID, "seed buffer too small");
      for (uint64_t r = 0; r < n_reads; r++) {
        uint64_t p0 = (read_off[r] - read_off[0]) + r;
        for (uint64_t i = key_off[r]; i < key_off[r + 1]; i++) {
          if (keys) keys[i] = kk[i];
          if (positions) positions[i] = start[i] - p0;
        }
      }
      return 0;
    }
    if (b->explicit_keys) {
      // key_win holds the window start; position = window start + delta
      std::vector<uint64_t> win(b->n_keys), kk(b->n_keys);
      std::vector<uint8_t> delta(b->positions, 0);  // long k-mers: keys sit at their start
      if (b->n_keys) {
        check_hip(hipMemcpyAsync(win.data(), b->key_win.p, b->n_keys * 8, hipMemcpyDeviceToHost, s), "copy");
        check_hip(hipMemcpyAsync(kk.data(), b->keys.p, b->n_keys * 8, hipMemcpyDeviceToHost, s), "copy");
        if (!b->keys_at_position)
          check_hip(hipMemcpyAsync(delta.data(), b->delta.p, b->positions, hipMemcpyDeviceToHost, s), "copy");
      }
      check_hip(hipStreamSynchronize(s), "copy");
      for (uint64_t r = 0; r <= n_reads; r++) key_off[r] = b->h_seg_off[r];
      if (b->n_keys > cap) fail(ZG_ERR_INVALID, "seed buffer too small");
      // position space (window start + selected offset) -> offset in its read
      for (uint64_t r = 0; r < n_reads; r++) {
        uint64_t p0 = (read_off[r] - read_off[0]) + r;
        for (uint64_t i = key_off[r]; i < key_off[r + 1]; i++) {
          if (keys) keys[i] = kk[i];
          if (positions) positions[i] = win[i] + delta[win[i]] - p0;
        }
      }
      return 0;
    }
    // k-mers: every valid window in order
    std::vector<uint64_t> codes((b->positions + 31) / 32 + 4), inv((b->positions + 63) / 64 + 4);
    check_hip(hipMemcpyAsync(codes.data(), b->codes.p, codes.size() * 8, hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipMemcpyAsync(inv.data(), b->invalid.p, inv.size() * 8, hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipStreamSynchronize(s), "copy");
    // the looked-up windows: k-mers, or findere's s-mers
    const uint32_t k = (uint32_t)(idx->m.index_mode == 3 ? idx->m.k - idx->m.findere_z : idx->m.k);
    auto bits = [](const std::vector<uint64_t> &w, uint64_t bit) {
      uint64_t i = bit >> 6, sh = bit & 63;
      return sh ? (w[i] >> sh) | (w[i + 1] << (64 - sh)) : w[i];
    };
    uint64_t n = 0;
    const uint64_t wmask = k >= 64 ? ~0ULL : ((1ULL << k) - 1);
    for (uint64_t r = 0; r < n_reads; r++) {
      key_off[r] = n;
      uint64_t p0 = (read_off[r] - read_off[0]) + r, len = read_off[r + 1] - read_off[r];
      for (uint64_t p = p0; p < p0 + len; p++) {
        if (bits(inv, p) & wmask) continue;
        uint64_t key = canonical_from_window(bits(codes, 2 * p), (int)k);
        if (idx->spaced.n) key = spaced_canonical(key, idx->spaced);
        if (idx->m.index_mode == 2 && mixsplit(key, idx->m.seed) % idx->m.modimizer_sampling)
          continue;  // not a modimizer
        if (n < cap) {
          if (keys) keys[n] = key;
          if (positions) positions[n] = p - p0;
        }
        n++;
      }
    }
    key_off[n_reads] = n;
    if (n > cap) fail(ZG_ERR_INVALID, "seed buffer too small");
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

// ---------------------------------------------------------------------------
// CLI-equivalent commands
// ---------------------------------------------------------------------------
namespace zg {

struct FdWriter {
  int fd;
  std::string buf;
  explicit FdWriter(int f) : fd(f) {}
  ~FdWriter() { flush(); }
  void put(const std::string &s) {
    buf += s;
    if (buf.size() > (1 << 20)) flush();
  }
  void flush() {
    size_t off = 0;
    while (fd >= 0 && off < buf.size()) {
      ssize_t w = write(fd, buf.data() + off, buf.size() - off);
      if (w <= 0) break;
      off += (size_t)w;
    }
    buf.clear();
  }
};

static std::string fmt_ratio(double r) {
  char b[64];
  snprintf(b, sizeof b, "%.8f", r);
  return b;
}

static std::string fmt_secs(double s) {
  char b[64];
  snprintf(b, sizeof b, "%.6f", s);
  return b;
}

static const char *mode_label(uint64_t m) {
  switch (m) {
    case 0: return "kmers";
    case 1: return "minimizers";
    case 2: return "modimizers";
    default: return "findere";
  }
}

// TSV field quoting (record_output.rs:67-84)
static void write_field(std::string &out, const std::string &v) {
  if (v.find_first_of("\t\r\n\"") == std::string::npos) {
    out += v;
    return;
  }
  out += '"';
  for (char c : v) {
    out += c;
    if (c == '"') out += '"';
  }
  out += '"';
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace zg

namespace zg {
// The query file parsed on a thread of its own into batches of at most
// `per_batch` records, at most two batches ahead of the consumer
// (records.rs:58-148: the parser thread and its bounded channel).
class BatchStream {
 public:
  BatchStream(const std::string &path, bool strict, size_t per_batch)
      : lines_(open_lines(path)), reader_(*lines_, strict), per_batch_(per_batch) {
    thread_ = std::thread([this] { run(); });
  }
  ~BatchStream() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
  }
  // The next batch; false after the last one. Rethrows a parse error.
  bool next(Records &out) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !queue_.empty() || done_; });
    if (queue_.empty()) {
      if (error_) std::rethrow_exception(error_);
      return false;
    }
    out = std::move(queue_.front());
    queue_.pop_front();
    cv_.notify_all();
    return true;
  }
  uint64_t records() const { return reader_.records(); }
  double parse_s() const { return parse_s_; }

 private:
  void run() {
    try {
      for (;;) {
        const double t = now_s();
        Records r;
        const bool more = reader_.next(r, per_batch_);
        parse_s_ += now_s() - t;
        if (!more) break;
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return queue_.size() < 2 || stop_; });
        if (stop_) break;
        queue_.push_back(std::move(r));
        cv_.notify_all();
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      error_ = std::current_exception();
    }
    std::lock_guard<std::mutex> lk(mu_);
    done_ = true;
    cv_.notify_all();
  }
  static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
  std::unique_ptr<LineSource> lines_;
  RecordReader reader_;
  size_t per_batch_;
  std::thread thread_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Records> queue_;
  bool done_ = false, stop_ = false;
  std::exception_ptr error_;
  double parse_s_ = 0;
};
}  // namespace zg

extern "C" int zg_query_command(const char *dir, const char *query, double min_ratio,
                                int recover, int out_fd, int err_fd) {
  return zg_query_command_ex(dir, query, min_ratio, recover, ZG_MEMORY_ALL, 0, out_fd, err_fd);
}

extern "C" int zg_query_command_ex(const char *dir, const char *query, double min_ratio,
                                   int recover, uint32_t memory, uint64_t hbm_budget, int out_fd,
                                   int err_fd) {
  zg_index *idx = nullptr;
  try {
    ensure(dir && query, "null argument", ZG_ERR_INVALID);
    ensure(min_ratio >= 0.0 && min_ratio <= 1.0, "--min-ratio must be in [0, 1]", ZG_ERR_INVALID);
    double t0 = now_s();
    int rc = zg_index_open(dir, &idx);
    if (rc) return rc;
    std::unique_ptr<zg_index> hold(idx);
    check_query_support(idx->m);
    // sequences parsed in bounded batches on their own thread while the
    // index loads and earlier batches are looked up
    BatchStream batches(query, false, 1 << 16);
    double tl = now_s();
    rc = zg_index_load_ex(idx, recover, hbm_budget, memory);
    if (rc) return rc;
    double load = now_s() - tl + (tl - t0);
    const uint32_t C = (uint32_t)idx->m.colors.size();
    std::vector<uint64_t> totals, agg(C), part(C);
    uint64_t total = 0;
    double lookup = 0;
    Records r;
    while (batches.next(r)) {
      totals.assign(r.size(), 0);
      double tq = now_s();
      rc = zg_query_reads(idx, (const uint8_t *)r.seq.data(), r.off.data(), r.size(), 0,
                          totals.data(), nullptr, part.data());
      if (rc) return rc;
      lookup += now_s() - tq;
      for (uint64_t t : totals) total += t;
      for (uint32_t c = 0; c < C; c++) agg[c] += part[c];
    }
    const double prepare = batches.parse_s();
    const Manifest &m = idx->m;
    uint64_t classes = m.classes();
    double to = now_s();
    {
      FdWriter out(out_fd);
      std::string s = "color\tmatches\tratio\tunique_kmers\n";
      for (uint32_t c = 0; c < C; c++) {
        double ratio = total == 0 ? 0.0 : (double)agg[c] / (double)total;
        if (ratio >= min_ratio)
          s += m.colors[c].name + "\t" + std::to_string(agg[c]) + "\t" + fmt_ratio(ratio) + "\t" +
               std::to_string(m.colors[c].unique_kmers) + "\n";
        if (s.size() > (1 << 20)) {
          out.put(s);
          s.clear();
        }
      }
      out.put(s);
    }
    double output = now_s() - to;
    FdWriter err(err_fd);
    err.put(std::string("#query\t") + query + "\tk=" + std::to_string(m.k) + "\tindex_mode=" +
            mode_label(m.index_mode) + "\ttotal_features=" + std::to_string(total) +
            "\tstack_probes=" + std::to_string(total) + "\tbank_probes=" +
            std::to_string(total * classes) + "\telapsed_s=" + fmt_secs(now_s() - t0 - load) + "\n");
    err.put("query_memory\tmode=gpu-resident\tmax_loaded_subindexes=" +
            std::to_string(m.banks.size()) + "\tsubindexes=" + std::to_string(m.banks.size()) +
            "\tresult_storage=HBM\n");
    err.put("query_preparation\tinput_passes=1\tseed_extraction_passes=1\tabandoned_file_scans=" +
            std::to_string(idx->recover ? 1 : 0) + "\tquery_features=" + std::to_string(total) +
            "\tqueried_recovery_hashes=" + std::to_string(idx->recover ? idx->rec.size() : 0) +
            "\tlookup=gpu-persistent\n");
    err.put("query_phases\tprepare_s=" + fmt_secs(prepare) + "\trecovery_map_s=0.000000\tlookup_s=" +
            fmt_secs(lookup) + "\toutput_s=" + fmt_secs(output) + "\n");
    err.put("query_timings\tload_s=" + fmt_secs(load) + "\tquery_s=" +
            fmt_secs(now_s() - t0 - load) + "\n");
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_query_records_command(const char *dir, const char *query, double min_ratio,
                                        int recover, uint64_t batch_size, int format, int out_fd,
                                        int err_fd) {
  return zg_query_records_command_ex(dir, query, min_ratio, recover, batch_size, format,
                                     ZG_MEMORY_ALL, 0, out_fd, err_fd);
}

extern "C" int zg_query_records_command_ex(const char *dir, const char *query, double min_ratio,
                                           int recover, uint64_t batch_size, int format,
                                           uint32_t memory, uint64_t hbm_budget, int out_fd,
                                           int err_fd) {
  zg_index *idx = nullptr;
  try {
    ensure(dir && query, "null argument", ZG_ERR_INVALID);
    ensure(min_ratio >= 0.0 && min_ratio <= 1.0, "--min-ratio must be in [0, 1]", ZG_ERR_INVALID);
    ensure(batch_size > 0, "record batch size must be positive", ZG_ERR_INVALID);
    ensure(format == 0 || format == 1, "invalid output format", ZG_ERR_INVALID);
    double t0 = now_s();
    int rc = zg_index_open(dir, &idx);
    if (rc) return rc;
    std::unique_ptr<zg_index> hold(idx);
    check_query_support(idx->m);
    // records parsed in batches of batch_size on their own thread, while the
    // index loads and the previous batch is looked up and written
    BatchStream batches(query, true, (size_t)std::min<uint64_t>(batch_size, SIZE_MAX));
    rc = zg_index_load_ex(idx, recover, hbm_budget, memory);
    if (rc) return rc;
    double load = now_s() - t0;
    const Manifest &m = idx->m;
    const uint32_t C = (uint32_t)m.colors.size();
    FdWriter out(out_fd);
    // bounded chunks of reads (per-read counters are n_reads x colours)
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(1ULL << 20, (1ULL << 28) / std::max(1u, C)));
    std::vector<uint64_t> totals;
    std::vector<uint32_t> counts;
    bool header = false;
    Records r;
    while (batches.next(r)) {
    if (!header) {
      out.put(format == 0 ? "query_id\tdataset_id\tdataset_name\tmatches\tqueried_seeds\tratio\n"
                          : "query_id\tvalid_kmers\thits\n");
      header = true;
    }
    for (uint64_t r0 = 0; r0 < r.size(); r0 += chunk) {
      uint64_t n = std::min<uint64_t>(chunk, r.size() - r0);
      totals.assign(n, 0);
      counts.assign(n * C, 0);
      rc = zg_query_reads(idx, (const uint8_t *)r.seq.data(), r.off.data() + r0, n,
                          ZG_QUERY_RECORDS, totals.data(), counts.data(), nullptr);
      if (rc) return rc;
      std::string s;
      for (uint64_t i = 0; i < n; i++) {
        const std::string &name = r.names[r0 + i];
        uint64_t total = totals[i];
        const uint32_t *row = counts.data() + i * C;
        // hits: all colours at ratio 0, else those passing the threshold
        // (query_stream.rs:139-161; output sorted by dataset id)
        auto pass = [&](uint32_t c) {
          if (min_ratio == 0.0) return true;
          return total != 0 && (double)row[c] / (double)total >= min_ratio;
        };
        if (format == 1) {
          s += name + "\t" + std::to_string(total) + "\t";
          bool firsth = true;
          for (uint32_t c = 0; c < C; c++) {
            if (!pass(c)) continue;
            if (!firsth) s += ',';
            firsth = false;
            s += std::to_string(c) + ":" + std::to_string(row[c]);
          }
          s += '\n';
        } else {
          bool any = false;
          for (uint32_t c = 0; c < C; c++) {
            if (!pass(c)) continue;
            any = true;
            write_field(s, name);
            s += "\t" + std::to_string(c) + "\t";
            write_field(s, m.colors[c].name);
            double ratio = total == 0 ? 0.0 : (double)row[c] / (double)total;
            s += "\t" + std::to_string(row[c]) + "\t" + std::to_string(total) + "\t" +
                 fmt_ratio(ratio) + "\n";
          }
          if (!any) {
            write_field(s, name);
            s += "\t.\t.\t0\t" + std::to_string(total) + "\t0.00000000\n";
          }
        }
        if (s.size() > (1 << 20)) {
          out.put(s);
          s.clear();
        }
      }
      out.put(s);
    }
    }
    ensure(batches.records() > 0, "no FASTA/FASTQ records", -1);
    out.flush();
    FdWriter err(err_fd);
    err.put("query_memory\tmode=gpu-resident\tmax_loaded_subindexes=" +
            std::to_string(m.banks.size()) + "\tsubindexes=" + std::to_string(m.banks.size()) +
            "\tresult_storage=HBM\n");
    err.put("query_timings\tload_s=" + fmt_secs(load) + "\tquery_s=" + fmt_secs(now_s() - t0 - load) +
            "\n");
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_info_command(const char *dir, int out_fd) {
  try {
    ensure(dir, "null argument", ZG_ERR_INVALID);
    Manifest m = load_manifest(dir);
    RecoveryMap rec;
    if (m.version == 1) {
      // version 1 loads whole, its fallback keys becoming the recovery map
      // (storage.rs:3903-3904)
      zg_index v1;
      v1.m = m;
      v1.dir = dir;
      load_v1(&v1);
      rec = std::move(v1.rec);
    } else {
      // validate payloads one at a time (storage.rs:3905-3913)
      for (auto &b : m.banks) (void)load_bank_payload(dir, b, (int)host_threads());
      if (m.abandoned_recovery) rec = read_recovery_map(dir, m);
    }
    uint64_t groups = 0, logical = 0, stack = 0, abandoned_keys = 0;
    for (auto &b : m.banks) {
      groups += b.groups.size();
      logical += b.slots * b.row_bits;
      stack += b.payload_bytes();
      for (auto &g : b.groups) abandoned_keys += g.has_abandoned_keys ? g.abandoned_keys : g.fallback_keys;
    }
    uint64_t largest = 0, partitions = 0;
    for (auto &b : m.banks) {
      largest = std::max(largest, b.payload_bytes());
      partitions += b.routed;
    }
    // fallback and metadata bytes (storage.rs:3939-3955)
    uint64_t fallback = m.abandoned_hashes * 8 + m.abandoned_origins * 4;
    if (m.version == 1) {
      fallback = 0;
      for (auto &b : m.banks)
        for (auto &g : b.groups) fallback += g.fallback_keys * 16;
    }
    const std::string manifest_path = path_join(dir, "manifest.json");
    struct stat st {};
    const uint64_t manifest_bytes = stat(manifest_path.c_str(), &st) == 0 ? (uint64_t)st.st_size : 0;
    const uint64_t metadata_bytes = manifest_bytes + m.banks.size() * 40 +
                                    (m.version == 1 ? groups * 8 : m.has_abandoned_file ? 8 : 0);
    uint64_t disk_bytes = 0;
    if (DIR *d = opendir(dir)) {
      while (struct dirent *e = readdir(d)) {
        const std::string n = e->d_name;
        if (n == "." || n == "..") continue;
        if (stat(path_join(dir, n).c_str(), &st) == 0) disk_bytes += (uint64_t)st.st_size;
      }
      closedir(d);
    }
    // optional manifest fields with the reference's serde defaults
    auto num = [&](const char *k, const char *dflt) -> std::string {
      const Json *v = m.json.get(k);
      return v && v->kind == Json::Number ? v->text : dflt;
    };
    auto secs = [&](const char *k) {
      const Json *v = m.json.get(k);
      char buf[64];
      snprintf(buf, sizeof buf, "%.6f", v && v->kind == Json::Number ? strtod(v->text.c_str(), nullptr) : 0.0);
      return std::string(buf);
    };
    auto str = [&](const char *k, const char *dflt) -> std::string {
      const Json *v = m.json.get(k);
      return v && v->kind == Json::String ? v->text : dflt;
    };
    auto flag = [&](const char *k, bool dflt) {
      const Json *v = m.json.get(k);
      return (v && v->kind == Json::Bool ? v->b : dflt) ? "true" : "false";
    };
    const std::string cells = str("construction_cell_storage", "in-memory-u64");
    const bool copied = cells == "copied-compressed-banks";
    std::string s;
    auto line = [&](const std::string &k, const std::string &v) { s += k + "\t" + v + "\n"; };
    size_t docs = std::max<size_t>(m.colors.size(), m.json.get("input_document_count")
                                                        ? strtoull(m.json.get("input_document_count")->text.c_str(), nullptr, 10)
                                                        : 0);
    const Json *skipped = m.json.get("skipped_documents");
    const size_t n_skipped = skipped && skipped->kind == Json::Array ? skipped->items.size() : 0;
    // info_command (storage.rs:3901-4090), line for line
    line("collection_status", n_skipped == 0 && docs == m.colors.size() ? "complete" : "incomplete");
    line("indexed_documents", std::to_string(m.colors.size()));
    if (copied) {
      std::set<uint64_t> comps;
      for (auto &b : m.banks) comps.insert(b.component);
      line("components", std::to_string(comps.size()));
    }
    line("index", dir);
    line("k", std::to_string(m.k));
    line("hashes", std::to_string(m.hashes));
    line("seed", std::to_string(m.seed));
    line("fingerprint_bits", std::to_string(m.fingerprint_bits));
    line("index_mode", mode_label(m.index_mode));
    line("technical_layout", "merge-pow2");
    line("colors", std::to_string(m.colors.size()));
    line("groups", std::to_string(groups));
    line("banks", std::to_string(m.banks.size()));
    line("size_classes", std::to_string(m.classes()));
    line("max_subindex_size", m.has_max_subindex_size ? std::to_string(m.max_subindex_size) : "unlimited");
    line("largest_subindex_bytes", std::to_string(largest));
    line("hash_partitions", std::to_string(partitions));
    line("input_documents", std::to_string(docs));
    line("skipped_documents", std::to_string(n_skipped));
    line("memory_guard_split_groups", num("memory_guard_split_groups", "0"));
    if (m.index_mode == 3) {
      line("findere_z", std::to_string(m.findere_z));
      line("indexed_k", std::to_string(m.k - m.findere_z));
    }
    line("minimizer_size", std::to_string(m.minimizer_size));
    line("modimizer_sampling", std::to_string(m.modimizer_sampling));
    line("minimizer_scheme", m.minimizer_scheme ? "open-closed-mod" : "simd");
    line("homopolymer_compression", m.homopolymer_compression ? "true" : "false");
    // anchors::print_settings (src/anchors.rs:112-126)
    line("seed_pattern", m.has_seed_pattern ? m.seed_pattern : "none");
    line("randstrobemers", m.has_randstrobemers ? "true" : "false");
    if (m.has_randstrobemers) {
      line("strobe_min", std::to_string(m.strobe_min));
      line("strobe_max", std::to_string(m.strobe_max));
    }
    line("merge_planner", copied ? "preserved" : m.merge_planner_label());
    line("planning_pair_evaluations", num("planning_pair_evaluations", "0"));
    line("planning_possible_pair_evaluations", num("planning_possible_pair_evaluations", "0"));
    line("merge_reverts_enabled", flag("merge_reverts_enabled", true));
    line("construction_cell_storage", cells);
    if (const Json *v = m.json.get("merge_memory_budget_bytes"); v && v->kind == Json::Number)
      line("merge_memory_budget_bytes", v->text);
    line("construction_abandoned_storage", str("construction_abandoned_storage", "legacy-memory"));
    if (m.has_segment_length) line("segment_length", std::to_string(m.segment_length));
    line("overflow_groups", num("overflow_groups", "0"));
    line("merge_column_order", copied ? "preserved" : str("column_order", "width"));
    if (const Json *v = m.json.get("column_order_method"); v && v->kind == Json::String)
      line("column_order_method", v->text);
    line("abandoned_recovery", m.abandoned_recovery ? "true" : "false");
    line("abandoned_hash_bits", "32");
    line("abandoned_keys", std::to_string(abandoned_keys));
    line("abandoned_unique_hashes", std::to_string(rec.size()));
    line("abandoned_dataset_associations", std::to_string(rec.origins.size()));
    // this library's recovery directory (open addressing, 8 B entries) and
    // postings; the reference prints its own OriginMap's heap bytes here
    uint64_t cap = 4;
    while (cap < (4 * rec.size() + 2) / 3) cap <<= 1;
    line("abandoned_map_ram_bytes",
         std::to_string(rec.size() ? cap * 8 + (rec.size() + rec.origins.size()) * 8 : 0));
    line("logical_bits", std::to_string(logical));
    line("padding_bits", std::to_string(stack * 8 - logical));
    line("stack_bytes", std::to_string(stack));
    line("stack_stored_bytes", std::to_string(stack));
    line("fallback_bytes", std::to_string(fallback));
    line("structure_bytes", std::to_string(stack + fallback));
    line("singleton_logical_bits", num("singleton_logical_bits", "0"));
    line("predicted_logical_bits", num("predicted_logical_bits", "0"));
    line("planned_groups", num("planned_groups", "0"));
    line("rejected_groups", num("rejected_groups", "0"));
    line("exact_group_verifications", num("exact_group_verifications", "0"));
    line("sketch_seconds", secs("sketch_seconds"));
    line("plan_seconds", secs("plan_seconds"));
    line("build_seconds", secs("build_seconds"));
    line("metadata_bytes", std::to_string(metadata_bytes));
    line("uncompressed_index_bytes", std::to_string(stack + fallback + metadata_bytes));
    line("disk_bytes", std::to_string(disk_bytes));
    line("min_subindex_documents", num("min_subindex_documents", "1"));
    line("promoted_groups", num("promoted_groups", "0"));
    if (m.has_max_subindexes) line("max_subindexes", std::to_string(m.max_subindexes));
    s += "bank\tslots\tgroups\trow_bits\tpadding_bits_per_row\tfallback_keys\tdocuments\n";
    for (auto &b : m.banks) {
      uint64_t fk = 0, docs_b = 0;
      for (auto &g : b.groups) {
        fk += g.fallback_keys;
        docs_b += g.members.size();
      }
      s += b.file + "\t" + std::to_string(b.slots) + "\t" + std::to_string(b.groups.size()) + "\t" +
           std::to_string(b.row_bits) + "\t" + std::to_string(b.row_bytes() * 8 - b.row_bits) + "\t" +
           std::to_string(fk) + "\t" + std::to_string(docs_b) + "\n";
    }
    s += "group\tslots\twidth\tunion_cardinality\tfallback_keys\tcolors\n";
    uint64_t id = 0;
    for (auto &b : m.banks)
      for (auto &g : b.groups) {
        s += std::to_string(id++) + "\t" + std::to_string(b.slots) + "\t" + std::to_string(g.width) +
             "\t" + std::to_string(g.union_cardinality) + "\t" + std::to_string(g.fallback_keys) + "\t";
        for (size_t j = 0; j < g.members.size(); j++) {
          if (j) s += ",";
          s += m.colors[g.members[j]].name;
        }
        s += "\n";
      }
    FdWriter out(out_fd);
    out.put(s);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}
