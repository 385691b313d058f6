// zg_serve.hip -- a request server with continuous batching over a resident
// index. Requests queue in arrival order; each slot (a worker thread with its
// own HIP stream and batch buffers) takes the queued requests that fit it as
// soon as its previous batch is answered, runs the batch pipeline
// (H2D -> pack -> seeds -> persistent lookup per bank -> per-request sums ->
// D2H into the callers' buffers) and answers them. With two or more slots the
// next batch is copied, seeded and launched while the previous one still runs.
//
// A request's counts are per read: the lookup writes counter cells of the
// batch's own reads only (the read index space of a batch is the
// concatenation of its requests), so what a request shares a batch with never
// changes its result.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "zg_internal.h"

namespace zg {

void set_error(const std::string &msg);  // zg_api.hip

using Clock = std::chrono::steady_clock;

struct ServeRequest {
  uint64_t ticket = 0;
  const uint8_t *seq = nullptr;
  const uint64_t *read_off = nullptr;
  uint64_t n_reads = 0, bases = 0;
  uint64_t *totals = nullptr;
  uint32_t *counts = nullptr;
  uint64_t *aggregate = nullptr;
  Clock::time_point t_submit, t_admit, t_done;
  uint64_t batch_id = 0;
  uint32_t batch_requests = 0;
  int status = 1;  // 1 while queued or running
  std::string error;
  uint64_t seeds = 0;
};
using RequestPtr = std::shared_ptr<ServeRequest>;

// Per-batch choice of the lookup order. The read-ordered lookup moves H rows
// per seed from HBM; the sorted sweep pays the bank's segment windows once
// per batch (about half of H banks from HBM, the rest from L2) and writes and
// reads back one hit row per seed. Rates measured on MI355X: row gathers
// 6.5 TB/s, window fills 8 TB/s, hit-row writes 2 TB/s, sequential
// reads and the sort's passes ~4-5 TB/s, ~1 ms of launches per chunk.
struct SweepModel {
  bool on = false;
  int force = -1;                 // ZG_SERVE_SWEEP: 0 never, 1 every batch, -1 the model
  uint64_t positions = 0;         // chunk capacity of a slot
  double read_per_seed = 0;       // s per seed, read order
  double sweep_per_seed = 0;      // s per seed, sweep
  double sweep_fixed = 0;         // s per chunk, sweep
  bool use(uint64_t seeds) const {
    if (!on || !seeds || force == 0) return false;
    if (force == 1) return true;
    const uint64_t chunks = (seeds + positions - 1) / positions;
    return seeds * sweep_per_seed + chunks * sweep_fixed < seeds * read_per_seed;
  }
};

struct ServeSlot {
  zg_batch b;                     // buffers at slot capacity, own stream
  DevBuf req_agg;                 // per-request colour sums
  std::vector<uint64_t> h_roff, h_poff, h_totals;
  std::thread thread;
  ~ServeSlot() {
    hipStream_t s = b.stream;
    b.stream = nullptr;
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

static double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace zg

struct zg_server {
  zg_index *idx = nullptr;
  zg_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<zg::RequestPtr> pending;
  std::unordered_map<uint64_t, zg::RequestPtr> live;  // submitted, not yet waited for
  uint64_t next_ticket = 1, next_batch = 1;
  bool stopping = false;
  std::vector<std::unique_ptr<zg::ServeSlot>> slots;
  uint64_t device_bytes = 0;
  // Subindex mode (banks beyond the HBM budget streamed in per query): the
  // index has one streaming buffer, so one batch at a time streams its banks
  // and runs their lookups (QueryMemory::Subindex, mod.rs:52-59).
  std::mutex stream_mu;
  zg::SweepModel sweep;  // batches large enough run as the sorted sweep
  // statistics (under mu)
  uint64_t done_requests = 0, batches = 0, reads = 0, seeds = 0, sweep_batches = 0;
  uint32_t max_batch_requests = 0;
  double lookup_ms = 0;
  void worker(zg::ServeSlot *slot);
  // Runs a batch; returns whether it ran as the sorted sweep.
  bool run_batch(zg::ServeSlot *slot, const std::vector<zg::RequestPtr> &batch);
};

using namespace zg;

// Device bytes of one slot at the configured capacity.
static uint64_t slot_bytes(const zg_index *idx, const zg_server_config &c) {
  const uint64_t C = idx->m.colors.size();
  const uint64_t pos = c.slot_bases + c.slot_reads;
  uint64_t n = (c.slot_bases + 64) + 2 * (c.slot_reads + 1) * 8;  // seq, offsets
  n += (code_words(pos, idx->m.k) + invalid_words(pos, idx->m.k)) * 8;  // codes, invalid bits
  n += (c.slot_reads + 1) * 8;                                    // totals
  const uint64_t strands = idx->m.has_randstrobemers ? 2 : 1;     // both orientations
  n += std::max<uint64_t>(1, strands * c.slot_reads * C) * 4;     // counters
  n += C * 8 + 8 + (uint64_t)c.max_requests * C * 8;              // sums
  n += (idx->dbanks.size() + 1) * 4;                              // work counters
  if (idx->m.index_mode == 1) {
    const uint64_t blocks = (pos + ocm_block_positions() - 1) / ocm_block_positions();
    n += 2 * (pos + 64) + (blocks + 1) * 12 + 2 * (pos + 64) * 8 + (c.slot_reads + 1) * 8;
    n += scan_temp_bytes_u32_to_u64(blocks + 1);
  }
  if (idx->m.homopolymer_compression) {
    const uint64_t words = (pos + 63) / 64;
    n += (words + 1) * 20 + scan_temp_bytes_u32_to_u64(words + 1);
  }
  if (idx->m.index_mode == 3) n += std::max<uint64_t>(1, pos * idx->fd_row_words) * 4;
  if ((idx->m.index_mode == 0 || idx->m.index_mode == 2) && idx->m.k > 32) {
    const uint64_t words = (pos + 63) / 64;
    n += (words + 1) * 20 + (pos + 64) * 8 * 3 + (c.slot_reads + 1) * 8;
    n += scan_temp_bytes_u32_to_u64(words + 1);
  }
  if (idx->m.has_randstrobemers) {
    const uint64_t words = (pos + 63) / 64;
    n += (words + 1) * 28 + (pos + 64) * 8 * 5 + (2 * c.slot_reads + 1) * 8;
    n += scan_temp_bytes_u32_to_u64(words + 1);
  }
  return n;
}

static void init_slot(zg_index *idx, const zg_server_config &c, ServeSlot &sl) {
  zg_batch &b = sl.b;
  b.idx = idx;
  b.flags = ZG_QUERY_RECORDS | (c.flags & ZG_QUERY_NO_RECOVERY);
  b.records = true;
  check_hip(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking), "hipStreamCreate");
  for (auto &e : b.ev) check_hip(hipEventCreate(&e), "hipEventCreate");
  const uint64_t C = idx->m.colors.size();
  const uint64_t pos = c.slot_bases + c.slot_reads;
  b.seq.alloc(c.slot_bases + 64);
  b.read_off.alloc((c.slot_reads + 1) * 8);
  b.pos_off.alloc((c.slot_reads + 1) * 8);
  b.codes.alloc(code_words(pos, idx->m.k) * 8);
  b.invalid.alloc(invalid_words(pos, idx->m.k) * 8);
  b.totals.alloc((c.slot_reads + 1) * 8);
  b.counts.alloc(std::max<uint64_t>(1, (idx->m.has_randstrobemers ? 2 : 1) * c.slot_reads * C) * 4);
  b.agg.alloc(C * 8 + 8);
  b.queue.alloc((idx->dbanks.size() + 1) * 4);
  sl.req_agg.alloc(std::max<uint64_t>(1, (uint64_t)c.max_requests * C) * 8);
  if (idx->m.index_mode == 1) {
    // the OCM buffers at capacity, so batches never allocate
    const uint64_t blocks = (pos + ocm_block_positions() - 1) / ocm_block_positions();
    b.flag.alloc(pos + 64);
    b.delta.alloc(pos + 64);
    b.block_counts.alloc((blocks + 1) * 4);
    b.block_off.alloc((blocks + 1) * 8);
    b.keys.alloc((pos + 64) * 8);
    b.key_win.alloc((pos + 64) * 8);
    b.key_off.alloc((c.slot_reads + 1) * 8);
    b.scan_temp.alloc(scan_temp_bytes_u32_to_u64(blocks + 1));
  }
  if (idx->m.index_mode == 3) b.fd_hits.alloc(std::max<uint64_t>(1, pos * idx->fd_row_words) * 4);
  if ((idx->m.index_mode == 0 || idx->m.index_mode == 2) && idx->m.k > 32) {
    // long k-mers: per-position keys, compacted
    const uint64_t words = (pos + 63) / 64;
    b.rs_cand.reserve((words + 1) * 8);
    b.rs_counts.reserve((words + 1) * 4);
    b.rs_off.reserve((words + 1) * 8);
    b.rs_hash.reserve((pos + 64) * 8);
    b.keys.reserve((pos + 64) * 8);
    b.key_win.reserve((pos + 64) * 8);
    b.key_off.reserve((c.slot_reads + 1) * 8);
    b.scan_temp.reserve(scan_temp_bytes_u32_to_u64(words + 1));
  }
  if (idx->m.has_randstrobemers) {
    // candidates (at most one per position) and both orientations' keys
    const uint64_t words = (pos + 63) / 64;
    b.rs_cand.reserve((words + 1) * 8);
    b.rs_counts.reserve((words + 1) * 4);
    b.rs_off.reserve((words + 1) * 8);
    b.rs_inv_off.reserve((words + 1) * 8);
    b.key_win.reserve((pos + 64) * 8);
    b.rs_hash.reserve((pos + 64) * 8);
    b.rs_run.reserve((pos + 64) * 8);
    b.keys.reserve((2 * pos + 64) * 8);
    b.key_off.reserve((2 * c.slot_reads + 1) * 8);
    b.scan_temp.reserve(scan_temp_bytes_u32_to_u64(words + 1));
  }
  if (idx->m.homopolymer_compression) {
    const uint64_t words = (pos + 63) / 64;
    b.hpc_keep.alloc((words + 1) * 8);
    b.hpc_counts.alloc((words + 1) * 4);
    b.hpc_off.alloc((words + 1) * 8);
    b.scan_temp.reserve(scan_temp_bytes_u32_to_u64(words + 1));
  }
  sl.h_roff.reserve(c.slot_reads + 1);
  sl.h_poff.reserve(c.slot_reads + 1);
  sl.h_totals.reserve(c.slot_reads + 1);
}

bool zg_server::run_batch(ServeSlot *slot, const std::vector<RequestPtr> &batch) {
  zg_batch &b = slot->b;
  hipStream_t s = b.stream;
  const uint32_t C = (uint32_t)idx->m.colors.size();
  uint64_t n_reads = 0, bases = 0;
  for (auto &r : batch) {
    n_reads += r->n_reads;
    bases += r->bases;
  }
  b.n_reads = n_reads;
  b.bases = bases;
  b.positions = bases + n_reads;
  b.explicit_keys = false;
  b.n_keys = 0;
  if (n_reads == 0) {
    for (auto &r : batch)
      if (r->aggregate) memset(r->aggregate, 0, C * 8ULL);
    return false;
  }
  // offsets of the batch's concatenated reads (one separator position per read)
  auto &roff = slot->h_roff, &poff = slot->h_poff;
  roff.assign(n_reads + 1, 0);
  poff.assign(n_reads + 1, 0);
  uint64_t r_at = 0, b_at = 0;
  for (auto &r : batch) {
    const uint64_t r0 = r->n_reads ? r->read_off[0] : 0;
    for (uint64_t i = 0; i < r->n_reads; i++) {
      roff[r_at + i + 1] = b_at + (r->read_off[i + 1] - r0);
      poff[r_at + i + 1] = roff[r_at + i + 1] + r_at + i + 1;
    }
    if (r->bases)
      check_hip(hipMemcpyAsync(b.seq.as<uint8_t>() + b_at, r->seq + r0, r->bases,
                               hipMemcpyHostToDevice, s),
                "hipMemcpyAsync");
    r_at += r->n_reads;
    b_at += r->bases;
  }
  check_hip(hipMemcpyAsync(b.read_off.p, roff.data(), (n_reads + 1) * 8, hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
  check_hip(hipMemcpyAsync(b.pos_off.p, poff.data(), (n_reads + 1) * 8, hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
  seed_batch(idx, &b);
  configure(idx, &b);
  check_hip(hipEventRecord(b.ev[0], s), "hipEventRecord");
  enqueue_zero_outputs(idx, &b);
  check_hip(hipEventRecord(b.ev[1], s), "hipEventRecord");
  const bool by_sweep = sweep.use(b.positions);
  if (by_sweep) {
    // the seeds in bank order (zg_sweep.hip) into the slot's reserved chunk
    // buffers: per-read counts as the read-ordered lookup gives them
    b.h_seg_off = poff;
    sweep_run(idx, &b);
    check_hip(hipEventRecord(b.ev[2], s), "hipEventRecord");
  } else if (any_streamed(idx)) {
    // the streamed banks pass through the shared buffer: this batch's
    // lookups end before another batch streams a bank into it
    std::lock_guard<std::mutex> lk(stream_mu);
    enqueue_lookups(idx, &b);
    check_hip(hipEventRecord(b.ev[2], s), "hipEventRecord");
    check_hip(hipEventSynchronize(b.ev[2]), "serve lookups");
  } else {
    enqueue_lookups(idx, &b);
    check_hip(hipEventRecord(b.ev[2], s), "hipEventRecord");
  }
  // per-request results straight into the callers' buffers
  launch_fill_u64(slot->req_agg.as<uint64_t>(), (uint64_t)batch.size() * C, 0, s);
  slot->h_totals.assign(n_reads + 1, 0);
  if (n_reads)
    check_hip(hipMemcpyAsync(slot->h_totals.data(), b.totals.p, n_reads * 8, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
  r_at = 0;
  for (size_t j = 0; j < batch.size(); j++) {
    const ServeRequest &r = *batch[j];
    const uint32_t *cnt = b.counts.as<uint32_t>() + r_at * C;
    if (r.counts && r.n_reads && C)
      check_hip(hipMemcpyAsync(r.counts, cnt, r.n_reads * C * 4ULL, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    if (r.aggregate && C) {
      uint64_t *agg = slot->req_agg.as<uint64_t>() + j * C;
      launch_colsum(cnt, r.n_reads, C, agg, s);
      check_hip(hipMemcpyAsync(r.aggregate, agg, C * 8ULL, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    }
    r_at += r.n_reads;
  }
  check_hip(hipEventRecord(b.ev[3], s), "hipEventRecord");
  check_hip(hipStreamSynchronize(s), "serve batch");
  b.last_lookup_ms = elapsed_ms(b.ev[1], b.ev[2]);
  r_at = 0;
  for (auto &r : batch) {
    uint64_t seeds_r = 0;
    for (uint64_t i = 0; i < r->n_reads; i++) seeds_r += slot->h_totals[r_at + i];
    if (r->totals && r->n_reads) memcpy(r->totals, slot->h_totals.data() + r_at, r->n_reads * 8);
    r->seeds = seeds_r;
    r_at += r->n_reads;
  }
  return by_sweep;
}

void zg_server::worker(ServeSlot *slot) {
  check_hip(hipSetDevice(idx->device), "hipSetDevice");
  for (;;) {
    std::vector<RequestPtr> batch;
    uint64_t id = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stopping || !pending.empty(); });
      if (pending.empty()) return;  // stopping, and every request answered
      uint64_t n = 0, nb = 0;
      while (!pending.empty()) {
        const ServeRequest &r = *pending.front();
        if (!batch.empty() && (n + r.n_reads > cfg.slot_reads || nb + r.bases > cfg.slot_bases ||
                               batch.size() >= cfg.max_requests))
          break;
        n += r.n_reads;
        nb += r.bases;
        batch.push_back(pending.front());
        pending.pop_front();
      }
      id = next_batch++;
    }
    const Clock::time_point t_admit = Clock::now();
    for (auto &r : batch) {
      r->t_admit = t_admit;
      r->batch_id = id;
      r->batch_requests = (uint32_t)batch.size();
    }
    int status = 0;
    std::string error;
    bool by_sweep = false;
    try {
      by_sweep = run_batch(slot, batch);
    } catch (...) {
      status = translate_exception();
      error = zg_last_error();
    }
    const Clock::time_point t_done = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto &r : batch) {
        r->t_done = t_done;
        r->status = status;
        r->error = error;
        if (!status) {
          done_requests++;
          reads += r->n_reads;
          seeds += r->seeds;
        }
      }
      batches++;
      if (by_sweep) sweep_batches++;
      max_batch_requests = std::max(max_batch_requests, (uint32_t)batch.size());
      if (!status && slot->b.last_lookup_ms > 0) lookup_ms += slot->b.last_lookup_ms;
    }
    cv_done.notify_all();
  }
}

extern "C" void zg_server_config_default(zg_server_config *c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->slots = 2;
  c->max_requests = 256;
  c->slot_reads = 32768;
  c->slot_bases = 64ULL << 20;
}

extern "C" int zg_server_create(zg_index *idx, const zg_server_config *cfg, zg_server **out) {
  try {
    ensure(idx && out, "null argument", ZG_ERR_INVALID);
    ensure(idx->on_device, "index is not resident on the device (call zg_index_load)",
           ZG_ERR_INVALID);
    zg_server_config c;
    zg_server_config_default(&c);
    if (cfg) c = *cfg;
    ensure(c.slots >= 1 && c.slots <= 16, "server slots must be in 1..=16", ZG_ERR_INVALID);
    ensure(c.max_requests >= 1, "max_requests must be positive", ZG_ERR_INVALID);
    ensure(c.slot_reads >= 1 && c.slot_bases >= 1, "slot capacity must be positive",
           ZG_ERR_INVALID);
    ensure(c.slot_reads * (uint64_t)std::max<size_t>(1, idx->m.colors.size()) < (1ULL << 40),
           "slot_reads x colours is too large", ZG_ERR_INVALID);
    ensure(!(c.flags & ~(uint32_t)ZG_QUERY_NO_RECOVERY), "unknown server flags", ZG_ERR_INVALID);
    auto srv = std::make_unique<zg_server>();
    srv->idx = idx;
    srv->cfg = c;
    check_hip(hipSetDevice(idx->device), "hipSetDevice");
    uint64_t per_slot = slot_bytes(idx, c);
    // the sorted sweep for large batches: k-mer seeds (k <= 32), banks
    // resident, the index tuned to sweep; each slot reserves its chunk
    // buffers here (at most kServeHitBytes of hit rows), so batches never
    // allocate
    const Manifest &m = idx->m;
    if (idx->tuning.sweep && !any_streamed(idx) && sweep_layout_ok(idx) &&
        (m.index_mode == 0 || m.index_mode == 2) && m.k <= 32 && !m.has_randstrobemers) {
      constexpr uint64_t kServeHitBytes = 16ULL << 30;
      const uint32_t rw = sweep_row_words(idx);  // builds the tables and bit planes
      SweepModel &sm = srv->sweep;
      sm.positions = std::max<uint64_t>(1, std::min<uint64_t>(c.slot_bases + c.slot_reads,
                                                              kServeHitBytes / (rw * 4ULL)));
      double gathered = 0, windows = 0;
      for (size_t i = 0; i < idx->dbanks.size(); i++) {
        const BankDev &v = idx->dbanks[i].view;
        gathered += (double)v.hashes * v.row_bytes / std::max<uint32_t>(1, v.partitions);
        windows += (double)v.hashes * idx->dbanks[i].data.bytes;
      }
      sm.read_per_seed = gathered / 6.5e12;
      sm.sweep_per_seed = rw * 4.0 / 2.0e12 + rw * 4.0 / 5.0e12 + 84.0 / 4.5e12;
      sm.sweep_fixed = windows / 2 / 8.0e12 + 1e-3;
      sm.on = true;
      if (const char *e = getenv("ZG_SERVE_SWEEP")) sm.force = strcmp(e, "0") == 0 ? 0 : 1;
      per_slot += sweep_chunk_bytes(idx, sm.positions);
    }
    srv->device_bytes = resident_bytes(idx) + per_slot * c.slots;
    if (c.hbm_budget && srv->device_bytes > c.hbm_budget)
      fail(ZG_ERR_BUDGET, "the resident index (" + std::to_string(resident_bytes(idx)) + " B) and " +
                              std::to_string(c.slots) + " slots of " + std::to_string(per_slot) +
                              " B exceed the HBM budget of " + std::to_string(c.hbm_budget) + " B");
    for (uint32_t i = 0; i < c.slots; i++) {
      srv->slots.push_back(std::make_unique<ServeSlot>());
      init_slot(idx, c, *srv->slots.back());
      if (srv->sweep.on) {
        zg_batch &b = srv->slots.back()->b;
        b.sw_hit_budget = srv->sweep.positions * sweep_row_words(idx) * 4ULL;
        sweep_reserve(idx, &b, srv->sweep.positions);
      }
    }
    for (auto &sl : srv->slots) {
      ServeSlot *p = sl.get();
      zg_server *sp = srv.get();
      sl->thread = std::thread([sp, p] {
        try {
          sp->worker(p);
        } catch (...) {
          // hipSetDevice failed: nothing can run on this slot; the others serve
        }
      });
    }
    *out = srv.release();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_server_submit(zg_server *srv, const uint8_t *seq, const uint64_t *read_off,
                                uint64_t n_reads, uint64_t *totals, uint32_t *counts,
                                uint64_t *aggregate, uint64_t *ticket) {
  try {
    ensure(srv && ticket && (n_reads == 0 || read_off), "null argument", ZG_ERR_INVALID);
    auto r = std::make_shared<ServeRequest>();
    r->seq = seq;
    r->read_off = read_off;
    r->n_reads = n_reads;
    for (uint64_t i = 0; i < n_reads; i++)
      ensure(read_off[i + 1] >= read_off[i], "read offsets must be non-decreasing", ZG_ERR_INVALID);
    r->bases = n_reads ? read_off[n_reads] - read_off[0] : 0;
    ensure(r->bases == 0 || seq, "null sequence", ZG_ERR_INVALID);
    if (r->n_reads > srv->cfg.slot_reads || r->bases > srv->cfg.slot_bases)
      fail(ZG_ERR_BUDGET, "request of " + std::to_string(r->n_reads) + " reads and " +
                              std::to_string(r->bases) + " bases exceeds one server slot (" +
                              std::to_string(srv->cfg.slot_reads) + " reads, " +
                              std::to_string(srv->cfg.slot_bases) + " bases)");
    r->totals = totals;
    r->counts = counts;
    r->aggregate = aggregate;
    r->t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      ensure(!srv->stopping, "server is shutting down", ZG_ERR_INVALID);
      r->ticket = srv->next_ticket++;
      srv->live[r->ticket] = r;
      srv->pending.push_back(r);
    }
    srv->cv_work.notify_one();
    *ticket = r->ticket;
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

static void fill_stats(const ServeRequest &r, zg_request_stats *st) {
  if (!st) return;
  memset(st, 0, sizeof *st);
  st->ticket = r.ticket;
  st->batch_id = r.batch_id;
  st->batch_requests = r.batch_requests;
  st->status = r.status;
  st->n_reads = r.n_reads;
  st->seeds = r.seeds;
  st->queued_ms = ms_between(r.t_submit, r.t_admit);
  st->service_ms = ms_between(r.t_admit, r.t_done);
  st->latency_ms = ms_between(r.t_submit, r.t_done);
}

static int finish_request(zg_server *srv, std::unordered_map<uint64_t, RequestPtr>::iterator it,
                          zg_request_stats *st) {
  RequestPtr r = it->second;
  srv->live.erase(it);
  fill_stats(*r, st);
  if (r->status) set_error(r->error);
  return r->status;
}

extern "C" int zg_server_wait(zg_server *srv, uint64_t ticket, zg_request_stats *st) {
  try {
    ensure(srv, "null argument", ZG_ERR_INVALID);
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    ensure(it != srv->live.end(), "unknown or already released ticket", ZG_ERR_INVALID);
    RequestPtr r = it->second;
    srv->cv_done.wait(lk, [&] { return r->status != 1; });
    return finish_request(srv, srv->live.find(ticket), st);
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_server_poll(zg_server *srv, uint64_t ticket, zg_request_stats *st) {
  try {
    ensure(srv, "null argument", ZG_ERR_INVALID);
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    ensure(it != srv->live.end(), "unknown or already released ticket", ZG_ERR_INVALID);
    if (it->second->status == 1) return 0;
    int rc = finish_request(srv, it, st);
    return rc ? rc : 1;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_server_get_stats(zg_server *srv, zg_server_stats *o) {
  try {
    ensure(srv && o, "null argument", ZG_ERR_INVALID);
    std::lock_guard<std::mutex> lk(srv->mu);
    memset(o, 0, sizeof *o);
    o->requests = srv->done_requests;
    o->batches = srv->batches;
    o->reads = srv->reads;
    o->seeds = srv->seeds;
    o->max_batch_requests = srv->max_batch_requests;
    o->slots = srv->cfg.slots;
    o->lookup_ms = srv->lookup_ms;
    o->device_bytes = srv->device_bytes;
    o->sweep_batches = srv->sweep_batches;
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_server_destroy(zg_server *srv) {
  if (!srv) return 0;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stopping = true;
  }
  srv->cv_work.notify_all();
  for (auto &sl : srv->slots)
    if (sl->thread.joinable()) sl->thread.join();
  delete srv;
  return 0;
}

extern "C" int zg_host_alloc(uint64_t bytes, void **out) {
  try {
    ensure(out, "null argument", ZG_ERR_INVALID);
    *out = nullptr;
    check_hip(hipHostMalloc(out, bytes ? bytes : 8, hipHostMallocDefault), "hipHostMalloc");
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_host_free(void *p) {
  if (p) (void)hipHostFree(p);
  return 0;
}
