// zg_internal.h -- the objects behind the opaque handles of zindex_gpu.h.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zindex_gpu.h"
#include "zg_format.h"
#include "zg_kernels.h"

namespace zg {

// Device allocation: owned (freed on reset) or borrowed from memory the
// caller manages (never freed here).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), owned(o.owned) {
    o.p = nullptr;
    o.bytes = 0;
    o.owned = true;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      owned = o.owned;
      o.p = nullptr;
      o.bytes = 0;
      o.owned = true;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
  }
  // Views n bytes at q (not freed by this object).
  void borrow(void *q, size_t n) {
    reset();
    p = q;
    bytes = n;
    owned = false;
  }
  void alloc(size_t n);
  // Grows to at least n bytes (contents are not preserved when it grows).
  void reserve(size_t n) {
    if (!p || bytes < n) alloc(n);
  }
  template <class T> T *as() const { return (T *)p; }
};

void check_hip(hipError_t e, const char *what);

// A size class of the sorted sweep: the banks (partitions) sharing one
// geometry and one colour -> column map. One radix sort per chunk serves
// them all; their hits go to one region of the hit row.
struct SweepClass {
  std::vector<uint32_t> banks;
  uint64_t count = 0;            // first segments of the geometry
  uint32_t slice_bytes = 0;      // LDS row stride (plane bytes per slice, <= 512)
  uint32_t n_slices = 0;
  uint32_t col_word0 = 0;        // first word of its column bits in a row
  uint32_t member_bit0 = 0;      // first bit of its member bits in a row
  uint32_t members = 0;
  bool planes = false;           // looked up over bit-plane copies of its banks
};

struct DeviceBank {
  DevBuf data, col_color, mg_off, mg_members, mg_mask_bit, color_col, color_bit;
  BankDev view{};
  // the payload as bit planes for the sorted sweep (built on first use)
  DevBuf planes;
  uint64_t planes_pitch = 0;
};

// Index options: the persistent lookup's launch shape.
struct Tuning {
  uint32_t span = 4096;
  uint32_t tiles_per_wave = 1;
  uint32_t prefetch = 2;
  uint32_t waves_per_cu = 0;  // 0 = occupancy limit
  uint32_t nontemporal = 0;   // non-temporal row loads
  uint32_t sweep = 0;         // seeds in segment order (sorted sweep)
  uint32_t graph = 0;         // the sweep of a batch as one HIP graph
};

}  // namespace zg

struct zg_index {
  zg::Manifest m;
  mutable std::string manifest_text;             // serialized on first request
  std::string dir;                               // empty when built in memory
  std::vector<std::vector<uint8_t>> payload;     // host bank payloads
  bool payload_loaded = false;
  zg::RecoveryMap rec;
  bool have_rec = false;                         // recovery map present in host memory
  bool recover = true;                           // use it in queries
  // device state
  bool on_device = false;
  int device = -1;
  uint32_t cu_count = 0;
  hipStream_t stream = nullptr;
  std::vector<zg::DeviceBank> dbanks;
  // Bank residency: a bank either owns its HBM payload (dbanks[i].data) or,
  // in subindex mode beyond the budget, is streamed into `stream_data` right
  // before its lookups (one bank at a time).
  std::vector<bool> streamed;                   // per bank
  std::vector<std::vector<uint32_t>> perm;      // device column order per bank
  std::vector<bool> as_is;                      // payload is already the device layout
  zg::DevBuf stream_data;
  int stream_holds = -1;                        // bank in stream_data, -1 none
  void *staging[2] = {nullptr, nullptr};        // page-locked upload chunks
  hipEvent_t staging_ev[2] = {nullptr, nullptr};
  zg::DevBuf rec_dir, rec_post, rec_post_col, rec_filter, inner_table;
  zg::SpacedRuns spaced;  // the manifest's seed pattern (n = 0: none)
  // findere: per-seed hit rows. Banks with the same colour -> bit map (the
  // partitions of one size class) share a region of fd_row_words words.
  struct FdRegion {
    uint32_t off = 0, bits = 0;  // first word in a row, bits in use
    uint64_t color_at = 0;       // first entry of its bit -> colour table
  };
  uint64_t fd_row_words = 0;
  std::vector<FdRegion> fd_regions;
  std::vector<uint32_t> fd_bank_region, fd_bank_member_bit;
  zg::DevBuf fd_bit_color;
  // sorted sweep: size classes, hit-row words (column regions, then member
  // bits) and the colour of every row bit
  std::vector<zg::SweepClass> sw_classes;
  uint32_t sw_row_words = 0, sw_col_words = 0, sw_members = 0;
  zg::DevBuf sw_bit_color;
  int sw_planes = -1;  // whether the tables were built with bit-plane copies allowed
  std::mutex sw_mu;    // guards building the tables
  zg::RecDev rec_view{};
  zg::Tuning tuning;
  ~zg_index();
};

struct zg_batch {
  zg_index *idx = nullptr;
  hipStream_t stream = nullptr;     // the index's stream, or a server slot's own
  uint32_t flags = 0;
  bool records = false;
  uint64_t n_reads = 0, bases = 0, positions = 0;
  std::vector<uint64_t> h_seg_off;  // read starts in the seed index space (host copy)
  zg::DevBuf seq, read_off, pos_off, codes, invalid, totals;
  // explicit seeds (minimizer mode)
  bool explicit_keys = false;
  uint64_t n_keys = 0;
  zg::DevBuf flag, delta, block_counts, block_off, keys, key_win, key_off, scan_temp;
  // homopolymer compression: keep flags, counts and offsets per 64 positions
  zg::DevBuf hpc_keep, hpc_counts, hpc_off;
  zg::DevBuf fd_hits;  // findere: positions x fd_row_words hit words
  // sorted sweep (zg_sweep.hip), per chunk: entries (first hash, position)
  // of the two radix passes, the valid-window bitmap, tile histograms and
  // their scan, segment starts, per-XCD work counters (+ the fold count);
  // hit rows go to fd_hits
  zg::DevBuf sw_h[2], sw_p[2], sw_listed, sw_hist, sw_off, sw_start, sw_queue;
  std::vector<hipEvent_t> sw_ev;  // per chunk: sort, lookup and sum bounds
  uint64_t sweep_chunks = 0;      // chunks of the last sweep
  uint32_t sweep_classes = 0;     // its size classes (a sort and lookups each per chunk)
  double sweep_lookup_ms = 0, sweep_sort_ms = 0, sweep_sum_ms = 0;  // their phase times
  double sweep_recovery_ms = 0;
  uint64_t sweep_folded = 0;      // entries answered from the entry before them
  // the sweep as one graph (tuning.graph): captured after an eager run, with
  // that run's chunk size
  hipGraph_t sw_graph = nullptr;
  hipGraphExec_t sw_exec = nullptr;
  bool sw_graph_failed = false, sw_capturing = false;
  uint64_t sw_cap = 0;            // chunk capacity (positions) of the last eager sweep
  uint64_t sw_hit_budget = 0;     // hit-row bytes per chunk at most (0: the library default)
  // the last chunk's sorted seed positions (last class) and their device count
  const uint32_t *sw_last_p = nullptr;
  const uint64_t *sw_last_n = nullptr;
  uint64_t sw_last_items = 0;
  // randstrobemers: candidate bitmap, per-word counts and offsets, run-id
  // offsets, candidate hashes and run ids (starts go to key_win); keys hold
  // the forward keys then the reverse-complement keys (rs_forward each)
  zg::DevBuf rs_cand, rs_counts, rs_off, rs_inv_off, rs_hash, rs_run;
  uint64_t rs_forward = 0;
  bool keys_at_position = false;  // explicit keys sit at their window start (long k-mers)
  // outputs
  zg::DevBuf counts, partial, agg, queue;
  uint64_t waves = 0;
  uint32_t blocks = 0;
  zg::KernelConfig kcfg{};
  uint32_t span = 0, n_tg = 0;
  uint32_t nontemporal = 0;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_failed = false;         // capture refused: launch directly
  // run stages; ev[4], ev[5] bracket each sorted-sweep lookup launch
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double last_lookup_ms = 0, last_seed_ms = 0, last_total_ms = 0;
  ~zg_batch();
};

struct zg_builder;

// Host worker threads (hardware count capped by OMP_NUM_THREADS, default 16).
unsigned zg_host_threads();

namespace zg {
// Builds the device copy of idx (banks, column tables, recovery directory).
void upload_index(zg_index *idx, uint64_t budget, uint32_t memory);
// Device bytes held by the resident index.
uint64_t resident_bytes(const zg_index *idx);
// Makes bank i's payload readable by lookups enqueued next on stream s
// (streams it in when the bank is not resident).
void make_resident(zg_index *idx, size_t i, hipStream_t s);
// Some bank streams per query (subindex mode beyond the budget).
bool any_streamed(const zg_index *idx);
// Bank payload in host memory, decoded from the bank file on first use.
const std::vector<uint8_t> &host_payload(zg_index *idx, size_t bank);
// The sorted sweep (zg_sweep.hip): whether an index's layout allows it, the
// lookups of a batch enqueued with no host synchronization, and the phase
// times and fold count of the last run (after the stream has drained).
bool sweep_layout_ok(const zg_index *idx);
void sweep_run(zg_index *idx, zg_batch *b);
void sweep_times(zg_batch *b, double &sort_ms, double &lookup_ms, double &recovery_ms,
                 double &sum_ms);
uint64_t sweep_folded(zg_batch *b);
// The sweep's tables and bit-plane copies built ahead (thread-safe); device
// bytes of the chunk buffers for `positions` positions of k-mer seeds, and
// reserving them in a batch; the hit-row words per position.
void sweep_prepare(zg_index *idx);
uint64_t sweep_chunk_bytes(zg_index *idx, uint64_t positions);
void sweep_reserve(zg_index *idx, zg_batch *b, uint64_t positions);
uint32_t sweep_row_words(zg_index *idx);
// GB/s of writing one hit row per listed seed of the last sweep chunk, at its
// position, in the sweep's order, with nothing else (the lookup's stores alone).
double sweep_store_probe(zg_batch *b);
// Batch pipeline pieces, all enqueued on b->stream (zg_api.hip).
void configure(zg_index *idx, zg_batch *b);
void seed_batch(zg_index *idx, zg_batch *b);
void enqueue_zero_outputs(zg_index *idx, zg_batch *b);
void enqueue_lookups(zg_index *idx, zg_batch *b);
// The lookup arguments of one bank for a seeded batch.
LookupArgs lookup_args(zg_index *idx, zg_batch *b, size_t bank);
double elapsed_ms(hipEvent_t a, hipEvent_t b);
int translate_exception();
// Host-side helpers shared with the builder.
std::vector<uint64_t> ocm_inner_table(uint32_t s, uint64_t seed);

// Result of peeling one group's equations on the GPU (zg_peel.hip). With
// remaining == 0 the group is solved and `cells` holds its verified cell
// values; otherwise the 2-core is left and slots/count/xr/state/order carry
// the device state (state: 0 active, 1 peeled; order: peel order).
struct DevicePeel {
  uint64_t remaining = 0;
  uint32_t rounds = 0;
  std::vector<uint64_t> cells;
  std::vector<uint32_t> slots, count, xr, order_e, order_p;
  std::vector<uint8_t> state;
};
void device_peel(const std::vector<uint64_t> &hashed, const std::vector<uint64_t> &values,
                 uint64_t S, uint32_t H, const SegmentGeometry &geo, DevicePeel &out);
}  // namespace zg
