// zg_build.hip -- host-side index construction in the reference on-disk format.
//
// Follows power_merge::build (src/power_merge/storage.rs:500-882) with a fixed
// plan: every dataset is its own group (or `merge_group_size` consecutive
// datasets share one), groups of equal slot count form one bank, columns are
// ordered by width then first member (ColumnOrder::Width, storage.rs:1644-1660),
// cells are found by ZOR hypergraph peeling (src/bin/zorindex.rs:634-841) with
// back-substitution (src/power_merge/graph.rs:163-216), and keys the peel
// abandons go to the recovery map (storage.rs:1311-1343, 2053-2081).
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

#include "zg_fastx.h"
#include "zg_hash.h"
#include "zg_internal.h"

namespace zg {

void set_error(const std::string &msg);  // zg_api.hip
int translate_exception();               // zg_api.hip

// ---------------------------------------------------------------------------
// Seed extraction on the host (build side).
// ---------------------------------------------------------------------------
static inline int enc(uint8_t c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}

// scan_long_kmers (src/lib.rs:683-730): rolling polynomial hashes of both
// orientations, the smaller one mixed with a k-dependent domain.
static void host_long_kmers(const uint8_t *seq, size_t len, int k, std::vector<uint64_t> &out) {
  const LongKmerParams lp = long_kmer_params((uint32_t)k);
  const uint64_t B = 0x9e3779b185ebca87ULL;
  uint64_t f = 0, r = 0, weight = 1;
  size_t seen = 0;
  for (size_t i = 0; i < len; i++) {
    int c = enc(seq[i]);
    if (c < 0) {
      f = r = 0;
      weight = 1;
      seen = 0;
      continue;
    }
    if (seen < (size_t)k) {
      f = f * B + (uint64_t)c + 1;
      r += (uint64_t)(4 - c) * weight;
      weight *= B;
      seen++;
    } else {
      const uint64_t o = (uint64_t)enc(seq[i - k]);
      f = (f - (o + 1) * lp.power) * B + (uint64_t)c + 1;
      r = (r - (4 - o)) * lp.inverse + (uint64_t)(4 - c) * lp.power;
    }
    if (seen == (size_t)k) out.push_back(splitmix64((f < r ? f : r) ^ lp.domain));
  }
}

// scan_canonical_kmers, src/lib.rs:645-681.
static void host_kmers(const uint8_t *seq, size_t len, int k, std::vector<uint64_t> &out) {
  if (k <= 0 || len < (size_t)k) return;
  if (k > 32) {
    host_long_kmers(seq, len, k, out);
    return;
  }
  uint64_t mask = k == 32 ? ~0ULL : ((1ULL << (2 * k)) - 1);
  int shift = 2 * (k - 1);
  uint64_t f = 0, r = 0;
  size_t seen = 0;
  for (size_t i = 0; i < len; i++) {
    int c = enc(seq[i]);
    if (c < 0) {
      f = r = 0;
      seen = 0;
      continue;
    }
    f = ((f << 2) | (uint64_t)c) & mask;
    r = (r >> 2) | ((uint64_t)(c ^ 3) << shift);
    if (++seen >= (size_t)k) out.push_back(f < r ? f : r);
  }
}

// Canonical rolling value of a fixed length (implementation.rs Rolling).
struct Rolling {
  uint64_t f = 0, r = 0, mask;
  int shift;
  explicit Rolling(int len) : mask(len >= 32 ? ~0ULL : ((1ULL << (2 * len)) - 1)), shift(2 * (len - 1)) {}
  uint64_t push(int c) {
    f = ((f << 2) | (uint64_t)c) & mask;
    r = (r >> 2) | ((uint64_t)(c ^ 3) << shift);
    return f < r ? f : r;
  }
};

// Sliding minimum with explicit window start (implementation.rs `Minimum`).
struct RankEntry {
  size_t pos;
  uint8_t cls;
  uint64_t hash;
};
struct SlidingMin {
  std::deque<RankEntry> q;
  bool rightmost;
  explicit SlidingMin(bool r) : rightmost(r) {}
  void push(size_t pos, uint8_t cls, uint64_t hash, size_t start) {
    while (!q.empty() && q.front().pos < start) q.pop_front();
    while (!q.empty()) {
      const RankEntry &b = q.back();
      bool greater = b.cls > cls || (b.cls == cls && b.hash > hash);
      bool equal = b.cls == cls && b.hash == hash;
      if (greater || (rightmost && equal)) q.pop_back();
      else break;
    }
    q.push_back({pos, cls, hash});
  }
};

// One all-ACGT run, the deque formulation of scan_run_scalar
// (implementation.rs:349-414).
static void host_ocm_run(const uint8_t *seq, size_t len, int k, int m, uint64_t seed,
                         std::vector<uint64_t> &out, std::vector<uint64_t> *positions = nullptr) {
  OcmParams P = ocm_parameters(k, m);
  const int w = P.w, t = P.t, s = P.s;
  const size_t inner_length = (size_t)(t - s + 1), anchor_length = (size_t)(k - t + 1);
  Rolling rep(m), anc(t), inner(s);
  SlidingMin inner_min(false), left(false), right(true);
  std::vector<uint64_t> inner_hashes(inner_length), reps((size_t)w);
  size_t gt = 0;
  bool have_prev = false;
  size_t prev = 0;
  for (size_t end = 0; end < len; end++) {
    int c = enc(seq[end]);
    gt += c >= 2;
    if (end >= (size_t)k) gt -= enc(seq[end - k]) >= 2;
    uint64_t key = rep.push(c), anchor_key = anc.push(c), inner_key = inner.push(c);
    if (end + 1 >= (size_t)m) reps[(end + 1 - m) % w] = key;
    if (end + 1 < (size_t)s) continue;
    size_t inner_pos = end + 1 - s;
    uint64_t h = mixsplit(inner_key, seed ^ kOcmInnerSeed);
    inner_hashes[inner_pos % inner_length] = h;
    size_t inner_start = end + 1 >= (size_t)t ? end + 1 - t : 0;
    inner_min.push(inner_pos, 0, h, inner_start);
    if (end + 1 < (size_t)t) continue;
    size_t apos = end + 1 - t;
    uint64_t mn = inner_min.q.front().hash;
    uint8_t cls;
    if (inner_hashes[(apos + (t - s) / 2) % inner_length] == mn) cls = 0;
    else if (inner_hashes[apos % inner_length] == mn || h == mn) cls = 1;
    else cls = 2;
    size_t start = apos + 1 >= anchor_length ? apos + 1 - anchor_length : 0;
    uint64_t rh = mixsplit(anchor_key, seed);
    left.push(apos, cls, rh, start);
    right.push(apos, cls, rh, start);
    if (end + 1 < (size_t)k) continue;
    size_t chosen = gt > (size_t)(k / 2) ? right.q.front().pos : left.q.front().pos;
    size_t pos = start + (chosen - start) % (size_t)w;
    if (!have_prev || prev != pos) {
      out.push_back(reps[pos % w]);
      if (positions) positions->push_back(pos);
      prev = pos;
      have_prev = true;
    }
  }
}

static void host_ocm(const uint8_t *seq, size_t len, int k, int m, uint64_t seed,
                     std::vector<uint64_t> &out) {
  if (k == 0 || (k & 1) == 0 || m == 0 || m > 31 || m > k) return;
  size_t start = 0;
  while (start < len) {
    size_t end = start;
    while (end < len && enc(seq[end]) >= 0) end++;
    if (end - start >= (size_t)k) host_ocm_run(seq + start, end - start, k, m, seed, out);
    start = end + 1;
  }
}

// The builder keeps every key as h = mixsplit(key, seed): splitmix64 is a
// bijection, h is PositionHash's `first` (storage.rs:3458-3465), and sorting
// by h both deduplicates keys and orders them by first segment (the segment
// is mulhi(h, count), monotone in h), which is the locality the peel wants.
static inline uint64_t key_of(uint64_t h, uint64_t seed) { return unsplitmix64(h) - seed; }

// keys -> sorted distinct h values: one MSD pass on the top 16 bits, then
// small sorts inside buckets (h is uniformly distributed).
static void hash_sort_unique(std::vector<uint64_t> &v, uint64_t seed) {
  for (uint64_t &x : v) x = mixsplit(x, seed);
  if (v.size() < 65536) {
    std::sort(v.begin(), v.end());
  } else {
    std::vector<uint64_t> tmp(v.size());
    std::vector<size_t> count(65537, 0);
    for (uint64_t x : v) count[(x >> 48) + 1]++;
    for (size_t i = 1; i < count.size(); i++) count[i] += count[i - 1];
    std::vector<size_t> cur(count.begin(), count.end() - 1);
    for (uint64_t x : v) tmp[cur[x >> 48]++] = x;
    for (size_t b = 0; b < 65536; b++) std::sort(tmp.begin() + count[b], tmp.begin() + count[b + 1]);
    v.swap(tmp);
  }
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// ---------------------------------------------------------------------------
// ZOR construction for one group.
// ---------------------------------------------------------------------------
struct GroupResult {
  std::vector<uint32_t> members;  // global colour ids
  uint64_t slots = 0;
  uint32_t width = 0;
  uint64_t union_cardinality = 0;
  uint64_t abandoned = 0;
  std::vector<uint8_t> cells;     // slots cells of cell_bytes each (little endian)
  uint32_t cell_bytes = 1;
  uint64_t cell(uint64_t s) const {
    uint64_t v = 0;
    memcpy(&v, cells.data() + s * cell_bytes, cell_bytes);
    return v;
  }
  // recovery records of abandoned keys: (hash32, member mask)
  std::vector<std::pair<uint32_t, uint64_t>> recovery;
  // hash partition of an oversized size class (partitions > 1)
  uint64_t source_slots = 0;
  uint32_t partition = 0, partitions = 1;
};

// A dataset's sequence kept at 2 bits per base plus an invalid-base bit, so
// its keys can be extracted again when its size class is hash-partitioned.
struct PackedColor {
  std::vector<uint64_t> codes, invalid, rec_off;
  void pack(const uint8_t *seq, const uint64_t *off, uint64_t n_records) {
    const uint64_t n = n_records ? off[n_records] - off[0] : 0;
    codes.assign((n + 31) / 32, 0);
    invalid.assign((n + 63) / 64, 0);
    rec_off.assign(off, off + n_records + 1);
    for (uint64_t i = 0; i < n; i++) {
      int c = enc(seq[off[0] + i]);
      if (c < 0) invalid[i >> 6] |= 1ULL << (i & 63);
      else codes[i >> 5] |= (uint64_t)c << (2 * (i & 31));
    }
  }
  // ACGT for valid bases, N for the rest (extraction treats both alike).
  std::vector<uint8_t> unpack() const {
    const uint64_t n = rec_off.empty() ? 0 : rec_off.back() - rec_off.front();
    std::vector<uint8_t> out(n);
    for (uint64_t i = 0; i < n; i++)
      out[i] = (invalid[i >> 6] >> (i & 63)) & 1 ? 'N' : "ACGT"[(codes[i >> 5] >> (2 * (i & 31))) & 3];
    return out;
  }
};

static uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// BuildConfig::slots (src/power_merge/mod.rs:319-337).
static uint64_t config_slots(const zg_build_config &c, uint64_t cardinality) {
  double needed = std::ceil(std::max<double>((double)cardinality, 1.0) * c.slot_scale);
  uint64_t n = (uint64_t)needed;
  uint64_t h2 = next_pow2(c.hashes);
  n = std::max<uint64_t>(n, h2);
  n = std::max<uint64_t>(n, (uint64_t)std::max<uint32_t>(c.segment_length, 1) * h2);
  return next_pow2(n);
}

// Set of cell ids with "largest member" queries: a 64-ary bit tree, so insert,
// erase and max each touch one word per level.
class MaxSet {
 public:
  explicit MaxSet(uint64_t n) {
    do {
      n = (n + 63) / 64;
      lv_.emplace_back(n, 0);
    } while (n > 1);
  }
  void insert(uint32_t x) {
    uint64_t v = x;
    for (auto &l : lv_) {
      uint64_t &w = l[v >> 6];
      bool had = w != 0;
      w |= 1ULL << (v & 63);
      if (had) return;
      v >>= 6;
    }
  }
  void erase(uint32_t x) {
    uint64_t v = x;
    for (auto &l : lv_) {
      uint64_t &w = l[v >> 6];
      w &= ~(1ULL << (v & 63));
      if (w) return;
      v >>= 6;
    }
  }
  bool empty() const { return lv_.back()[0] == 0; }
  uint32_t max() const {  // requires !empty()
    uint64_t v = 0;
    for (size_t l = lv_.size(); l-- > 0;) v = v * 64 + 63 - __builtin_clzll(lv_[l][v]);
    return (uint32_t)v;
  }

 private:
  std::vector<std::vector<uint64_t>> lv_;
};

// Peeling with degree counts and XOR-of-incident-edge accumulators; a cyclic
// core is broken at the minimum-degree cell (highest cell id on ties) by
// keeping its first active equation and abandoning the others, the default
// no-heuristic rule of choose_cycle_break_candidate. `hashed` holds the
// group's distinct keys as ascending h = mixsplit(key, seed), i.e. ordered by
// first segment, so the degree/XOR updates walk memory nearly sequentially;
// every key's H cells are computed once. Once peeling stalls, the cells of
// degree d >= 2 sit in bucket d (a MaxSet), and each cell's equations are
// listed in ascending id order (counting sort by cell).
static void construct_group(const zg_build_config &cfg, const std::vector<uint64_t> &hashed,
                            const std::vector<uint64_t> &values, GroupResult &g) {
  const uint64_t S = g.slots;
  const uint32_t H = cfg.hashes;
  const SegmentGeometry geo = segment_geometry(S, H, cfg.segment_length);
  const size_t n = hashed.size();
  ensure(S <= UINT32_MAX && n < UINT32_MAX, "group too large for 32-bit cell ids", -6);
  static const bool trace = getenv("ZG_BUILD_TRACE") != nullptr;
  std::chrono::steady_clock::time_point tp[8];
  auto T = [&](int i) { if (trace) tp[i] = std::chrono::steady_clock::now(); };
  T(0);
  std::vector<uint32_t> slots, count, xr;
  std::vector<uint8_t> state;  // 0 active, 1 peeled, 2 abandoned
  std::vector<std::pair<uint32_t, uint32_t>> order;
  size_t remaining = n;
  std::vector<uint64_t> cells;
  bool solved = false;  // cells final and verified (device peel without a core)
  if (cfg.device) {
    // parallel peeling rounds on the GPU; a 2-core comes back for the rule below
    DevicePeel dp;
    device_peel(hashed, values, S, H, geo, dp);
    if (trace) fprintf(stderr, "  device peel: %u rounds, %llu equations left\n", dp.rounds,
                       (unsigned long long)dp.remaining);
    if (dp.remaining == 0) {
      cells = std::move(dp.cells);
      solved = true;
      remaining = 0;
    } else {
      slots = std::move(dp.slots);
      count = std::move(dp.count);
      xr = std::move(dp.xr);
      state = std::move(dp.state);
      remaining = dp.remaining;
      order.reserve(n);
      for (size_t i = 0; i < dp.order_e.size(); i++) order.push_back({dp.order_e[i], dp.order_p[i]});
    }
    T(1);
  } else {
    // 1. the H cells of every equation (PositionHash::positions)
    slots.resize((size_t)n * H);
    for (size_t e = 0; e < n; e++) {
      uint64_t first = hashed[e], step = splitmix64(first);
      uint64_t base = slot_base(first, geo), h = first;
      for (uint32_t i = 0; i < H; i++) {
        slots[e * H + i] = (uint32_t)slot_at(base, h, i, geo);
        h += step;
      }
    }
    T(1);
    // 2. degrees and XOR of incident equation ids
    count.assign(S, 0);
    xr.assign(S, 0);
    for (size_t e = 0; e < n; e++)
      for (uint32_t i = 0; i < H; i++) {
        uint32_t s = slots[e * H + i];
        count[s]++;
        xr[s] ^= (uint32_t)e;
      }
    state.assign(n, 0);
    order.reserve(n);
  }
  T(2);
  std::vector<uint32_t> queue;
  if (!solved)
    for (uint64_t s = 0; s < S; s++)
      if (count[s] == 1) queue.push_back((uint32_t)s);
  size_t qi = 0;
  std::vector<uint32_t> abandoned;
  // core phase state
  bool core = false;
  std::vector<MaxSet> bucket;          // bucket[d]: cells of degree d >= 2
  std::vector<uint64_t> adj_off;       // cell -> [adj_off[s], adj_off[s+1]) in adj
  std::vector<uint32_t> adj;
  auto remove_edge = [&](uint32_t e) {
    for (uint32_t i = 0; i < H; i++) {
      uint32_t s = slots[(size_t)e * H + i];
      if (count[s] > 0) {
        if (core && count[s] > 1) bucket[count[s]].erase(s);
        count[s]--;
        xr[s] ^= e;
        if (count[s] == 1) queue.push_back(s);
        if (core && count[s] > 1) bucket[count[s]].insert(s);
      }
    }
  };
  while (remaining > 0) {
    bool progress = false;
    while (qi < queue.size()) {
      uint32_t s = queue[qi++];
      if (count[s] != 1) continue;
      uint32_t e = xr[s];
      if (state[e] != 0) continue;
      state[e] = 1;
      remaining--;
      progress = true;
      order.push_back({e, s});
      remove_edge(e);
    }
    if (qi > (1u << 20)) {
      queue.erase(queue.begin(), queue.begin() + qi);
      qi = 0;
    }
    if (remaining == 0) break;
    if (progress && qi < queue.size()) continue;
    if (!core) {
      core = true;
      if (trace) fprintf(stderr, "  core: %zu of %zu equations left after %.0f ms\n", remaining, n,
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp[2]).count());
      uint32_t dmax = 0;
      for (uint64_t s = 0; s < S; s++) dmax = std::max(dmax, count[s]);
      for (uint32_t d = 0; d <= dmax; d++) bucket.emplace_back(d >= 2 ? S : 1);
      for (uint64_t s = 0; s < S; s++)
        if (count[s] > 1) bucket[count[s]].insert((uint32_t)s);
      adj_off.assign(S + 1, 0);
      for (size_t e = 0; e < n; e++)
        if (state[e] == 0)
          for (uint32_t i = 0; i < H; i++) adj_off[slots[e * H + i] + 1]++;
      for (uint64_t s = 0; s < S; s++) adj_off[s + 1] += adj_off[s];
      adj.resize(adj_off[S]);
      std::vector<uint64_t> fill(adj_off.begin(), adj_off.end() - 1);
      for (size_t e = 0; e < n; e++)
        if (state[e] == 0)
          for (uint32_t i = 0; i < H; i++) adj[fill[slots[e * H + i]]++] = (uint32_t)e;
    }
    // the minimum-degree cell, highest id on ties
    uint32_t pick = UINT32_MAX;
    for (size_t d = 2; d < bucket.size(); d++)
      if (!bucket[d].empty()) {
        pick = bucket[d].max();
        break;
      }
    if (pick == UINT32_MAX) {
      // no multi-degree cell left: abandon the first active equation
      for (size_t e = 0; e < n; e++)
        if (state[e] == 0) {
          state[e] = 2;
          remaining--;
          abandoned.push_back((uint32_t)e);
          remove_edge((uint32_t)e);
          break;
        }
      continue;
    }
    bool kept = false;
    for (uint64_t j = adj_off[pick]; j < adj_off[pick + 1]; j++) {
      uint32_t e = adj[j];
      if (state[e] != 0) continue;
      if (!kept) {
        kept = true;
        continue;
      }
      state[e] = 2;
      remaining--;
      abandoned.push_back(e);
      remove_edge(e);
    }
  }
  T(3);
  std::vector<uint32_t>().swap(count);
  std::vector<uint32_t>().swap(xr);
  if (!solved) {
    // 3. back-substitution in reverse peel order (graph.rs:187-199)
    cells.assign(S, 0);
    for (size_t i = order.size(); i-- > 0;) {
      uint32_t e = order[i].first, pivot = order[i].second;
      uint64_t v = values[e];
      for (uint32_t h = 0; h < H; h++) {
        uint32_t s = slots[(size_t)e * H + h];
        if (s != pivot) v ^= cells[s];
      }
      cells[pivot] = v;
    }
  }
  T(4);
  // 4. verification (graph.rs:201-214); the device verified a solved group
  for (size_t e = 0; !solved && e < n; e++) {
    if (state[e] == 2) continue;
    uint64_t v = 0;
    for (uint32_t h = 0; h < H; h++) v ^= cells[slots[e * H + h]];
    if (v != values[e]) fail(-1, "incorrect merged payload for key " + std::to_string(key_of(hashed[e], cfg.seed)));
  }
  g.cell_bytes = g.width <= 8 ? 1 : g.width <= 16 ? 2 : g.width <= 32 ? 4 : 8;
  g.cells.resize(S * g.cell_bytes);
  if (g.cell_bytes == 1) {
    for (uint64_t s = 0; s < S; s++) g.cells[s] = (uint8_t)cells[s];
  } else {
    for (uint64_t s = 0; s < S; s++)
      memcpy(g.cells.data() + s * g.cell_bytes, &cells[s], g.cell_bytes);
  }
  T(5);
  g.abandoned = abandoned.size();
  std::sort(abandoned.begin(), abandoned.end());
  for (uint32_t e : abandoned) {
    uint64_t mask = g.members.size() == 1 ? 1 : values[e] >> cfg.fingerprint_bits;
    g.recovery.push_back({abandoned_hash(key_of(hashed[e], cfg.seed), cfg.seed), mask});
  }
  if (trace) {
    T(6);
    auto ms = [&](int a, int b) { return std::chrono::duration<double, std::milli>(tp[b] - tp[a]).count(); };
    fprintf(stderr, "  group: bucket %.0f, degrees %.0f, peel %.0f, assign %.0f, verify %.0f, pack %.0f ms\n",
            ms(0, 1), ms(1, 2), ms(2, 3), ms(3, 4), ms(4, 5), ms(5, 6));
  }
}

}  // namespace zg

// ---------------------------------------------------------------------------
// The builder object
// ---------------------------------------------------------------------------
struct zg_builder {
  zg_build_config cfg{};
  // deques: workers hold references to elements while the caller appends
  std::deque<zg::Color> colors;
  std::deque<std::vector<uint64_t>> pending;  // keys per colour
  std::deque<zg::PackedColor> packed;         // sequences, for re-extraction
  std::deque<zg::GroupResult> groups;         // finished groups (index = group id)
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> jobs;
  std::vector<std::thread> workers;
  size_t running = 0;
  bool stopping = false;
  std::string error;
  int error_code = 0;
  std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();

  void worker() {
    while (true) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stopping || !jobs.empty(); });
        if (jobs.empty()) return;
        job = std::move(jobs.front());
        jobs.pop_front();
        running++;
      }
      try {
        job();
      } catch (const zg::Error &e) {
        std::lock_guard<std::mutex> lk(mu);
        if (!error_code) { error = e.what(); error_code = e.code; }
      } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(mu);
        if (!error_code) { error = e.what(); error_code = -1; }
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        running--;
      }
      cv.notify_all();
    }
  }
  void submit(std::function<void()> f) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return jobs.size() < workers.size() * 2; });
    jobs.push_back(std::move(f));
    cv.notify_all();
  }
  void drain() {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return jobs.empty() && running == 0; });
    if (error_code) throw zg::Error(error_code, error);
  }
  ~zg_builder() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stopping = true;
    }
    cv.notify_all();
    for (auto &t : workers) t.join();
  }
};

namespace zg {

// homopolymer_compress (src/lib.rs:209-222): drops a base equal to the one
// before it (case-insensitive); other bytes are kept and break runs. Kept
// bases are written upper case. Returns the compressed length.
static size_t homopolymer_compress(const uint8_t *seq, size_t len, uint8_t *out) {
  size_t n = 0;
  int prev = -1;
  for (size_t i = 0; i < len; i++) {
    const int c = enc(seq[i]);
    if (c < 0 || c != prev) out[n++] = c < 0 ? seq[i] : (uint8_t)"ACGT"[c];
    prev = c;
  }
  return n;
}

static SpacedRuns config_spaced(const zg_build_config &cfg) {
  if (!cfg.seed_pattern[0]) return SpacedRuns{};
  uint32_t width = 0;
  uint64_t mask = 0;
  std::string err = parse_spaced_seed(cfg.seed_pattern, &width, &mask);
  ensure(err.empty(), err.c_str(), -1);
  return spaced_runs(width, mask);
}

// Linker (src/anchors.rs:185-256) over one run's candidates in position
// order: every anchor emits one key, paired with the later candidate at a
// distance in [min_d, max_d] of least rank (the earliest on ties), or with
// itself when there is none.
static void link_strobes(const std::vector<uint64_t> &pos, const std::vector<uint64_t> &hash,
                         uint64_t min_d, uint64_t max_d, std::vector<uint64_t> &out) {
  const size_t n = pos.size();
  for (size_t i = 0; i < n; i++) {
    uint64_t partner = hash[i];
    uint32_t best = 64;
    for (size_t j = i + 1; j < n; j++) {
      const uint64_t d = pos[j] - pos[i];
      if (d > max_d) break;
      if (d < min_d) continue;
      const uint32_t r = strobe_rank(hash[i], hash[j]);
      if (r < best) {
        best = r;
        partner = hash[j];
        if (r == 0) break;
      }
    }
    out.push_back(strobe_pair(hash[i], partner));
  }
}

static void strobe_bounds(const zg_build_config &cfg, uint64_t *lo, uint64_t *hi) {
  const uint64_t length = cfg.index_mode == 1 ? cfg.minimizer_size : cfg.k;
  *lo = cfg.strobe_min ? cfg.strobe_min : length;
  *hi = cfg.strobe_max ? cfg.strobe_max : 3 * length;
}

// anchors::scan with randstrobemers (src/anchors.rs:284-351) on one record:
// per run of valid bases, the canonical k-mers (or the selected minimizer
// positions, each once) become strobes, linked downstream.
static void extract_strobes(const zg_build_config &cfg, const SpacedRuns &spaced,
                            const uint8_t *p, size_t len, std::vector<uint64_t> &keys) {
  uint64_t lo = 0, hi = 0;
  strobe_bounds(cfg, &lo, &hi);
  std::vector<uint64_t> ckeys, cpos, chash;
  for (size_t a = 0; a < len;) {
    if (enc(p[a]) < 0) { a++; continue; }
    size_t b = a;
    while (b < len && enc(p[b]) >= 0) b++;
    ckeys.clear();
    cpos.clear();
    if (cfg.index_mode == 1) {
      if (b - a >= cfg.k) {
        host_ocm_run(p + a, b - a, (int)cfg.k, (int)cfg.minimizer_size, cfg.seed, ckeys, &cpos);
        std::vector<std::pair<uint64_t, uint64_t>> sel(cpos.size());
        for (size_t i = 0; i < cpos.size(); i++) sel[i] = {cpos[i], ckeys[i]};
        std::sort(sel.begin(), sel.end());
        sel.erase(std::unique(sel.begin(), sel.end(),
                              [](const std::pair<uint64_t, uint64_t> &x,
                                 const std::pair<uint64_t, uint64_t> &y) { return x.first == y.first; }),
                  sel.end());
        cpos.resize(sel.size());
        ckeys.resize(sel.size());
        for (size_t i = 0; i < sel.size(); i++) {
          cpos[i] = sel[i].first;
          ckeys[i] = sel[i].second;
        }
      }
    } else {
      host_kmers(p + a, b - a, (int)cfg.k, ckeys);
      cpos.resize(ckeys.size());
      for (size_t i = 0; i < cpos.size(); i++) cpos[i] = i;
    }
    chash.resize(ckeys.size());
    for (size_t i = 0; i < ckeys.size(); i++)
      chash[i] = strobe_hash(spaced.n ? spaced_canonical(ckeys[i], spaced) : ckeys[i]);
    link_strobes(cpos, chash, lo, hi, keys);
    a = b;
  }
}

// scan_index_features (src/lib.rs:573-642, anchors.rs:265-283) per record.
static void extract(const zg_build_config &cfg, const uint8_t *seq, const uint64_t *off,
                    uint64_t n, std::vector<uint64_t> &keys) {
  const SpacedRuns spaced = config_spaced(cfg);
  std::vector<uint8_t> compressed;
  for (uint64_t r = 0; r < n; r++) {
    const uint8_t *p = seq + off[r];
    size_t len = off[r + 1] - off[r];
    if (cfg.homopolymer_compression) {
      compressed.resize(len);
      len = homopolymer_compress(p, len, compressed.data());
      p = compressed.data();
    }
    const size_t at = keys.size();
    if (cfg.randstrobemers) {
      extract_strobes(cfg, spaced, p, len, keys);
      continue;
    }
    if (cfg.index_mode == 1) {
      host_ocm(p, len, (int)cfg.k, (int)cfg.minimizer_size, cfg.seed, keys);
    } else if (cfg.index_mode == 3) {
      // findere: canonical (k - z)-mers of the runs of at least k valid bases
      for (size_t a = 0; a < len;) {
        if (enc(p[a]) < 0) { a++; continue; }
        size_t b = a;
        while (b < len && enc(p[b]) >= 0) b++;
        if (b - a >= cfg.k) host_kmers(p + a, b - a, (int)(cfg.k - cfg.findere_z), keys);
        a = b;
      }
    } else {
      host_kmers(p, len, (int)cfg.k, keys);
      if (cfg.index_mode == 2) {  // scan_modimizers (src/lib.rs:796-806)
        size_t kept = at;
        for (size_t i = at; i < keys.size(); i++)
          if (mixsplit(keys[i], cfg.seed) % cfg.modimizer_sampling == 0) keys[kept++] = keys[i];
        keys.resize(kept);
      }
    }
    if (spaced.n)
      for (size_t i = at; i < keys.size(); i++) keys[i] = spaced_canonical(keys[i], spaced);
  }
}

// The group's equations: distinct keys (as h, ascending) and cell values
// fp | membership mask << b (storage.rs:1222-1255, 1301-1308). Consumes
// key_sets.
static void group_equations(const zg_build_config &cfg, const std::vector<uint32_t> &members,
                            std::vector<std::vector<uint64_t> *> key_sets,
                            std::vector<uint64_t> &keys, std::vector<uint64_t> &values) {
  const uint32_t b = cfg.fingerprint_bits;
  keys.clear();
  values.clear();
  if (members.size() == 1) {
    keys.swap(*key_sets[0]);
    values.resize(keys.size());
    for (size_t i = 0; i < keys.size(); i++) values[i] = fingerprint(key_of(keys[i], cfg.seed), cfg.seed, b);
  } else {
    // union with membership masks (storage.rs:1222-1255, 1301-1308)
    std::vector<std::pair<uint64_t, uint64_t>> all;
    for (size_t j = 0; j < key_sets.size(); j++)
      for (uint64_t key : *key_sets[j]) all.push_back({key, 1ULL << j});
    std::sort(all.begin(), all.end());
    for (size_t i = 0; i < all.size();) {
      uint64_t key = all[i].first, mask = 0;
      while (i < all.size() && all[i].first == key) mask |= all[i++].second;
      keys.push_back(key);
      values.push_back(fingerprint(key_of(key, cfg.seed), cfg.seed, b) | (mask << b));
    }
    for (auto *ks : key_sets) std::vector<uint64_t>().swap(*ks);
  }
}

static void build_group(const zg_build_config &cfg, std::vector<uint32_t> members,
                        std::vector<std::vector<uint64_t> *> key_sets, GroupResult &g) {
  g.members = std::move(members);
  std::vector<uint64_t> keys, values;
  group_equations(cfg, g.members, key_sets, keys, values);
  g.union_cardinality = keys.size();
  g.width = cfg.fingerprint_bits + (g.members.size() == 1 ? 0 : (uint32_t)g.members.size());
  g.slots = config_slots(cfg, keys.size());
  construct_group(cfg, keys, values, g);
}

}  // namespace zg

using namespace zg;

extern "C" void zg_build_config_default(zg_build_config *c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->k = 31;
  c->hashes = 4;
  c->fingerprint_bits = 8;
  c->index_mode = 0;
  c->minimizer_size = 21;
  c->segment_length = 64;
  c->seed = 69;
  c->slot_scale = 1.0;
  c->max_subindex_size = 8ULL << 30;
  c->merge_group_size = 1;
  c->threads = 0;
  c->drop_abandoned = 0;
  c->modimizer_sampling = 16;
  c->findere_z = 10;
}

static void validate_config(const zg_build_config &c) {
  uint32_t b = c.fingerprint_bits;
  ensure(b == 1 || b == 2 || b == 4 || b == 8 || b == 16 || b == 32,
         "fingerprint bits must be 1, 2, 4, 8, 16 or 32", -1);
  ensure(c.hashes >= 2 && c.hashes <= 32, "invalid hash count", -1);
  // 0 = no fixed segment length: the legacy geometry (manifest version 7)
  ensure(c.segment_length == 0 || (c.segment_length <= 4096 &&
                                   !(c.segment_length & (c.segment_length - 1))),
         "segment length must be 0 (legacy geometry) or a power of two in 1..=4096", -1);
  ensure(c.slot_scale > 0 && std::isfinite(c.slot_scale), "invalid slot scale", -1);
  ensure(c.max_subindex_size && !(c.max_subindex_size & (c.max_subindex_size - 1)),
         "--max-subindex-size must be a non-zero power of two", -1);
  if (c.index_mode == 0 || c.index_mode == 2) {
    ensure(c.index_mode == 0 || c.modimizer_sampling >= 1, "modimizer sampling must be positive",
           -1);
    ensure(c.k >= 1, "--k must be positive", -1);
    ensure(c.k <= 32 || !c.randstrobemers,
           "randstrobemers of k-mers longer than 32 are not built here", -6);
  } else if (c.index_mode == 1) {
    ensure(c.k >= 1 && c.k <= 255 && c.k % 2 == 1 && c.minimizer_size >= 1 &&
               c.minimizer_size <= 31 && c.minimizer_size <= c.k,
           "invalid minimizer window/length", -1);
  } else if (c.index_mode == 3) {
    // findere::indexed_length (src/findere.rs:17-21)
    ensure(c.k > 0, "findere query k must be positive", -1);
    ensure(c.findere_z < c.k, "--findere-z must be smaller than query k", -1);
    ensure(c.k - c.findere_z <= 32, "findere s-mers longer than 32 are not built here", -6);
  } else {
    fail(-1, "unsupported index mode");
  }
  if (c.seed_pattern[0] || c.randstrobemers)
    ensure(c.index_mode == 0 || c.index_mode == 1,
           "spaced seeds and randstrobemers require --index-mode kmers or minimizers", -1);
  if (c.seed_pattern[0]) {
    ensure(memchr(c.seed_pattern, 0, sizeof c.seed_pattern) != nullptr,
           "seed pattern width must be in 1..=31", -1);
    uint32_t width = 0;
    uint64_t mask = 0;
    const std::string err = parse_spaced_seed(c.seed_pattern, &width, &mask);
    ensure(err.empty(), err.c_str(), -1);
    const uint32_t length = c.index_mode == 1 ? c.minimizer_size : c.k;
    if (width != length)
      fail(-1, "seed pattern width " + std::to_string(width) + " must match " +
                   (c.index_mode == 1 ? "minimizer size m" : "k") + " (" +
                   std::to_string(length) + ")");
  }
  if (c.randstrobemers) {
    uint64_t lo = 0, hi = 0;
    strobe_bounds(c, &lo, &hi);
    // RandstrobeConfig::validate (src/anchors.rs:98-110)
    ensure(lo > 0, "randstrobemer minimum distance must be greater than zero", -1);
    ensure(lo <= hi, "randstrobemer minimum distance must not exceed maximum distance", -1);
  }
  ensure(c.merge_group_size >= 1 && c.merge_group_size <= 64 - b, "invalid merge group size",
         -1);
  ensure(c.device <= 1, "device must be 0 (CPU peel) or 1 (GPU peel)", -1);
}

extern "C" int zg_builder_create(const zg_build_config *cfg, zg_builder **out) {
  try {
    ensure(cfg && out, "null argument", -1);
    validate_config(*cfg);
    auto b = std::make_unique<zg_builder>();
    b->cfg = *cfg;
    uint32_t threads = cfg->threads ? cfg->threads : zg_host_threads();
    if (!threads) threads = 1;
    for (uint32_t i = 0; i < threads; i++) b->workers.emplace_back([p = b.get()] { p->worker(); });
    *out = b.release();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_builder_add_color(zg_builder *b, const char *name, const uint8_t *seq,
                                    const uint64_t *rec_off, uint64_t n_records) {
  try {
    ensure(b && name && (n_records == 0 || (seq && rec_off)), "null argument", -1);
    ensure(*name, "empty color name", -1);
    {
      std::lock_guard<std::mutex> lk(b->mu);
      if (b->error_code) throw Error(b->error_code, b->error);
    }
    uint32_t id;
    {
      std::lock_guard<std::mutex> lk(b->mu);
      id = (uint32_t)b->colors.size();
      Color c;
      c.name = name;
      c.path = name;
      b->colors.push_back(c);
    }
    uint64_t base = n_records ? rec_off[0] : 0;
    uint64_t len = n_records ? rec_off[n_records] - base : 0;
    auto copy = std::make_shared<std::vector<uint8_t>>(seq ? seq + base : nullptr,
                                                        seq ? seq + base + len : nullptr);
    auto offs = std::make_shared<std::vector<uint64_t>>(n_records + 1);
    for (uint64_t i = 0; i <= n_records; i++) (*offs)[i] = n_records ? rec_off[i] - base : 0;
    const zg_build_config cfg = b->cfg;
    const bool single = cfg.merge_group_size == 1;
    {
      std::lock_guard<std::mutex> lk(b->mu);
      b->pending.emplace_back();
      b->packed.emplace_back();
      if (single) b->groups.emplace_back();
    }
    b->submit([b, id, copy, offs, n_records, cfg, single] {
      static const bool trace = getenv("ZG_BUILD_TRACE") != nullptr;
      auto t0 = std::chrono::steady_clock::now();
      std::vector<uint64_t> keys;
      keys.reserve(copy->size());
      extract(cfg, copy->data(), offs->data(), n_records, keys);
      auto t1 = std::chrono::steady_clock::now();
      uint64_t occurrences = keys.size();
      hash_sort_unique(keys, cfg.seed);
      uint64_t unique = keys.size();
      auto t2 = std::chrono::steady_clock::now();
      std::vector<uint64_t> *slot;
      PackedColor *pk;
      GroupResult *gr = nullptr;
      {
        std::lock_guard<std::mutex> lk(b->mu);
        b->colors[id].occurrences = occurrences;
        b->colors[id].unique_kmers = unique;
        slot = &b->pending[id];
        pk = &b->packed[id];
        if (single) gr = &b->groups[id];
      }
      pk->pack(copy->data(), offs->data(), n_records);
      *slot = std::move(keys);
      if (single) build_group(cfg, {id}, {slot}, *gr);
      if (trace) {
        auto t3 = std::chrono::steady_clock::now();
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        fprintf(stderr, "build colour %u: extract %.1f ms, sort %.1f ms, group %.1f ms\n", id,
                ms(t0, t1), ms(t1, t2), ms(t2, t3));
      }
    });
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

namespace zg {
// partition_large_subindexes (storage.rs:1525-1641): a size class whose bank
// would exceed max_subindex_size is rebuilt as P = slots / local hash
// partitions of `local` slots (halved until the bank fits); each group's keys
// are extracted again from the kept sequences and split by partition_id.
static void partition_oversized(zg_builder *b) {
  const zg_build_config &cfg = b->cfg;
  std::map<uint64_t, uint64_t> widths;
  for (auto &g : b->groups) widths[g.slots] += g.width;
  const uint64_t minimum = (uint64_t)std::max<uint32_t>(cfg.segment_length, 1) * next_pow2(cfg.hashes);
  std::map<uint64_t, std::pair<uint64_t, uint32_t>> splits;
  for (auto &[slots, bits] : widths) {
    const uint64_t rb = (bits + 7) / 8;
    if ((unsigned __int128)slots * rb <= cfg.max_subindex_size) continue;
    uint64_t local = slots;
    while ((unsigned __int128)local * rb > cfg.max_subindex_size) local /= 2;
    if (local < minimum)
      fail(-1, "--max-subindex-size=" + std::to_string(cfg.max_subindex_size) +
                   " bytes is too small for " + std::to_string(rb) + " bytes per slice and minimum " +
                   std::to_string(minimum) + " slots; raise the cap");
    splits[slots] = {local, (uint32_t)(slots / local)};
  }
  if (splits.empty()) return;
  static const bool progress = getenv("ZG_BUILD_PROGRESS") != nullptr;
  std::atomic<size_t> rebuilt{0};
  size_t to_rebuild = 0;
  for (auto &g : b->groups) to_rebuild += splits.count(g.slots);
  if (progress)
    fprintf(stderr, "[build] hash partitioning %zu oversized size classes (%zu groups)\n",
            splits.size(), to_rebuild);
  std::deque<GroupResult> old;
  old.swap(b->groups);
  std::vector<std::vector<GroupResult>> parts(old.size());
  for (size_t gi = 0; gi < old.size(); gi++) {
    auto it = splits.find(old[gi].slots);
    if (it == splits.end()) continue;
    const uint64_t local = it->second.first;
    const uint32_t P = it->second.second;
    parts[gi].resize(P);
    b->submit([b, gi, local, P, &old, &parts, &rebuilt, to_rebuild] {
      const zg_build_config &c = b->cfg;
      GroupResult &src = old[gi];
      std::vector<std::vector<uint64_t>> sets(src.members.size());
      std::vector<std::vector<uint64_t> *> ptrs;
      for (size_t j = 0; j < src.members.size(); j++) {
        const PackedColor &pc = b->packed[src.members[j]];
        std::vector<uint8_t> seq = pc.unpack();
        std::vector<uint64_t> roff(pc.rec_off.size());
        for (size_t r = 0; r < roff.size(); r++) roff[r] = pc.rec_off[r] - pc.rec_off[0];
        extract(c, seq.data(), roff.data(), roff.empty() ? 0 : roff.size() - 1, sets[j]);
        hash_sort_unique(sets[j], c.seed);
        ptrs.push_back(&sets[j]);
      }
      std::vector<uint64_t> keys, values;
      group_equations(c, src.members, ptrs, keys, values);
      std::vector<std::vector<uint64_t>> pk(P), pv(P);
      for (size_t i = 0; i < keys.size(); i++) {
        uint32_t p = (uint32_t)partition_id(key_of(keys[i], c.seed), c.seed, P);
        pk[p].push_back(keys[i]);
        pv[p].push_back(values[i]);
      }
      for (uint32_t p = 0; p < P; p++) {
        GroupResult &g = parts[gi][p];
        g.members = src.members;
        g.width = src.width;
        g.slots = local;
        g.source_slots = src.slots;
        g.partition = p;
        g.partitions = P;
        g.union_cardinality = pk[p].size();
        construct_group(c, pk[p], pv[p], g);
        std::vector<uint64_t>().swap(pk[p]);
        std::vector<uint64_t>().swap(pv[p]);
      }
      std::vector<uint8_t>().swap(src.cells);
      size_t done = ++rebuilt;
      if (progress && (done % 64 == 0 || done == to_rebuild))
        fprintf(stderr, "[build] partitioned %zu/%zu groups into %u parts\n", done, to_rebuild, P);
    });
  }
  b->drain();
  for (size_t gi = 0; gi < old.size(); gi++) {
    if (parts[gi].empty()) {
      b->groups.push_back(std::move(old[gi]));
    } else {
      for (auto &g : parts[gi]) b->groups.push_back(std::move(g));
    }
  }
}

// Packs groups into plane-layout banks and assembles the manifest.
static void finish_index(zg_builder *b, zg_index *idx, double build_seconds) {
  const zg_build_config &cfg = b->cfg;
  const uint32_t fb = cfg.fingerprint_bits;
  // banks by (source slots, partition), ascending (pack_banks_streaming's
  // BTreeMap order, storage.rs:1762-1767)
  auto bank_key = [](const GroupResult &g) {
    return std::make_pair(g.partitions > 1 ? g.source_slots : g.slots, (uint64_t)g.partition);
  };
  std::vector<std::pair<uint64_t, uint64_t>> keys_of_banks;
  for (auto &g : b->groups) keys_of_banks.push_back(bank_key(g));
  std::sort(keys_of_banks.begin(), keys_of_banks.end());
  keys_of_banks.erase(std::unique(keys_of_banks.begin(), keys_of_banks.end()), keys_of_banks.end());
  Manifest &m = idx->m;
  m = Manifest();
  uint64_t singleton_bits = 0, logical_bits = 0, abandoned_total = 0;
  bool partitioned = false;
  for (auto bank_k : keys_of_banks) {
    std::vector<size_t> ids;
    for (size_t i = 0; i < b->groups.size(); i++)
      if (bank_key(b->groups[i]) == bank_k) ids.push_back(i);
    const GroupResult &g0 = b->groups[ids[0]];
    const uint64_t slots = g0.slots;
    std::stable_sort(ids.begin(), ids.end(), [&](size_t x, size_t y) {
      const GroupResult &a = b->groups[x], &c = b->groups[y];
      if (a.width != c.width) return a.width < c.width;
      return a.members[0] < c.members[0];
    });
    BankMeta bm;
    bm.slots = slots;
    bm.fingerprint_first = true;
    uint64_t row_bits = 0;
    for (size_t i : ids) {
      const GroupResult &g = b->groups[i];
      GroupMeta gm;
      gm.members = g.members;
      gm.width = g.width;
      gm.offset = row_bits;
      gm.union_cardinality = g.union_cardinality;
      gm.fallback_keys = cfg.drop_abandoned ? 0 : g.abandoned;
      gm.has_abandoned_keys = true;
      gm.abandoned_keys = g.abandoned;
      row_bits += g.width;
      bm.groups.push_back(gm);
      abandoned_total += g.abandoned;
    }
    bm.row_bits = row_bits;
    bm.routed = g0.partitions > 1;
    if (bm.routed) {
      partitioned = true;
      bm.routing_source_slots = g0.source_slots;
      bm.partition = g0.partition;
      bm.partitions = g0.partitions;
    }
    bm.file = BankMeta::filename(slots, bm.routed, g0.source_slots, g0.partition, g0.partitions);
    uint64_t rb = bm.row_bytes();
    ensure((unsigned __int128)slots * rb <= cfg.max_subindex_size,
           "internal: a bank exceeds --max-subindex-size after partitioning", -1);
    logical_bits += slots * row_bits;
    std::vector<uint8_t> payload(slots * rb, 0);
    const uint64_t G = ids.size();
    std::vector<uint64_t> mask_off(G);
    uint64_t acc = G * fb;
    for (size_t i = 0; i < G; i++) {
      mask_off[i] = acc;
      acc += b->groups[ids[i]].width - fb;
    }
    const uint64_t fp_mask = (1ULL << fb) - 1;
    const uint32_t threads = (uint32_t)b->workers.size();
    const uint64_t tile = 4096;
    const uint64_t tiles = (slots + tile - 1) / tile;
    std::atomic<uint64_t> next{0};
    auto work = [&] {
      for (uint64_t t; (t = next++) < tiles;) {
        uint64_t lo = t * tile, hi = std::min(slots, lo + tile);
        for (size_t i = 0; i < G; i++) {
          const GroupResult &g = b->groups[ids[i]];
          if (fb == 8 && g.width == 8) {
            const uint8_t *cells = g.cells.data();
            for (uint64_t s = lo; s < hi; s++) payload[s * rb + i] = cells[s];
            continue;
          }
          for (uint64_t s = lo; s < hi; s++) {
            uint8_t *row = payload.data() + s * rb;
            uint64_t v = g.cell(s);
            uint64_t fpv = v & fp_mask;
            uint64_t bit = i * fb;
            for (uint32_t x = 0; x < fb; x++)
              if ((fpv >> x) & 1) row[(bit + x) >> 3] |= (uint8_t)(1u << ((bit + x) & 7));
            if (g.width > fb) {
              uint64_t mk = v >> fb;
              uint64_t mb = mask_off[i];
              for (uint32_t x = 0; x < g.width - fb; x++)
                if ((mk >> x) & 1) row[(mb + x) >> 3] |= (uint8_t)(1u << ((mb + x) & 7));
            }
          }
        }
      }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    for (size_t i : ids) std::vector<uint8_t>().swap(b->groups[i].cells);
    idx->payload.push_back(std::move(payload));
    m.banks.push_back(std::move(bm));
  }
  for (auto &g : b->groups) singleton_bits += (uint64_t)fb * g.slots * g.members.size();
  // recovery map: union origins per hash (visit_abandoned)
  RecoveryMap rec;
  if (!cfg.drop_abandoned) {
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (auto &g : b->groups)
      for (auto &r : g.recovery) {
        uint64_t mask = r.second;
        while (mask) {
          int bit = __builtin_ctzll(mask);
          mask &= mask - 1;
          pairs.push_back({r.first, g.members[bit]});
        }
      }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    for (size_t i = 0; i < pairs.size();) {
      uint32_t h = pairs[i].first;
      while (i < pairs.size() && pairs[i].first == h) rec.origins.push_back(pairs[i++].second);
      rec.hashes.push_back(h);
      rec.off.push_back(rec.origins.size());
    }
  }
  // manifest fields (storage.rs:787-857)
  m.version = partitioned ? 9 : cfg.segment_length ? 8 : 7;
  m.abandoned_recovery = !cfg.drop_abandoned;
  m.has_abandoned_file = !cfg.drop_abandoned;
  m.abandoned_file = cfg.drop_abandoned ? "" : kAbandonedFile;
  m.abandoned_hashes = rec.hashes.size();
  m.abandoned_origins = rec.origins.size();
  m.backend = "zor";
  m.k = cfg.k;
  m.hashes = cfg.hashes;
  m.has_segment_length = cfg.segment_length != 0;
  m.segment_length = cfg.segment_length;
  m.seed = cfg.seed;
  m.fingerprint_bits = fb;
  m.index_mode = cfg.index_mode;
  m.minimizer_size = cfg.minimizer_size;
  m.minimizer_scheme = 1;
  m.modimizer_sampling = cfg.modimizer_sampling ? cfg.modimizer_sampling : 16;
  m.findere_z = cfg.findere_z;
  m.homopolymer_compression = cfg.homopolymer_compression != 0;
  if (cfg.randstrobemers) {
    m.has_randstrobemers = true;
    strobe_bounds(cfg, &m.strobe_min, &m.strobe_max);
    m.randstrobe_scheme = "strobealign-popcount8-symmetric64-downstream-v1";
  }
  if (cfg.seed_pattern[0]) {
    m.has_seed_pattern = true;
    m.seed_pattern = cfg.seed_pattern;
    parse_spaced_seed(m.seed_pattern, &m.seed_width, &m.seed_mask);
  }
  m.slot_scale = cfg.slot_scale;
  m.has_max_subindex_size = true;
  m.max_subindex_size = cfg.max_subindex_size;
  m.colors.assign(b->colors.begin(), b->colors.end());
  Json j = Json::object();
  j.set("version", Json::u64(m.version));
  j.set("abandoned_recovery", Json::boolean(m.abandoned_recovery));
  j.set("abandoned_file", m.has_abandoned_file ? Json::str(m.abandoned_file) : Json::null());
  j.set("abandoned_hashes", Json::u64(m.abandoned_hashes));
  j.set("abandoned_origins", Json::u64(m.abandoned_origins));
  j.set("backend", Json::str("zor"));
  j.set("k", Json::u64(m.k));
  j.set("hashes", Json::u64(m.hashes));
  j.set("segment_length", m.has_segment_length ? Json::u64(m.segment_length) : Json::null());
  j.set("seed", Json::u64(m.seed));
  j.set("fingerprint_bits", Json::u64(fb));
  j.set("index_mode", Json::u64(m.index_mode));
  j.set("minimizer_size", Json::u64(m.minimizer_size));
  j.set("minimizer_scheme", Json::str("open-closed-mod"));
  j.set("homopolymer_compression", Json::boolean(m.homopolymer_compression));
  j.set("seed_pattern", cfg.seed_pattern[0] ? Json::str(cfg.seed_pattern) : Json::null());
  if (m.has_randstrobemers) {
    Json r = Json::object();
    r.set("min_distance", Json::u64(m.strobe_min));
    r.set("max_distance", Json::u64(m.strobe_max));
    j.set("randstrobemers", r);
    j.set("randstrobe_scheme", Json::str(m.randstrobe_scheme));
  } else {
    j.set("randstrobemers", Json::null());
    j.set("randstrobe_scheme", Json::null());
  }
  j.set("modimizer_sampling", Json::u64(m.modimizer_sampling));
  j.set("findere_z", Json::u64(m.findere_z));
  j.set("sketch_size", Json::u64(128));
  j.set("sketch_method", Json::str("none: fixed consecutive grouping (no sketches)"));
  j.set("merge_planner", Json::str("naive"));
  j.set("column_order", Json::str("width"));
  j.set("column_order_method", Json::str("width-then-dataset-id-v1"));
  j.set("planning_pair_evaluations", Json::u64(0));
  j.set("planning_possible_pair_evaluations", Json::u64(0));
  j.set("merge_reverts_enabled", Json::boolean(false));
  j.set("construction_cell_storage", Json::str("two-pass-memory-packed-no-scratch"));
  j.set("construction_abandoned_storage", Json::str(cfg.drop_abandoned ? "omitted" : "memory"));
  j.set("slot_scale", Json::f64(cfg.slot_scale));
  j.set("min_subindex_documents", Json::u64(1));
  j.set("max_subindex_size", Json::u64(cfg.max_subindex_size));
  j.set("promoted_groups", Json::u64(0));
  j.set("singleton_logical_bits", Json::u64(singleton_bits));
  j.set("predicted_logical_bits", Json::u64(logical_bits));
  j.set("planned_groups", Json::u64(b->groups.size()));
  j.set("rejected_groups", Json::u64(0));
  j.set("overflow_groups", Json::u64(0));
  uint64_t verifications = b->groups.size();
  if (partitioned) {
    verifications = 0;
    for (auto &g : b->groups)
      verifications += g.union_cardinality - (cfg.drop_abandoned ? g.abandoned : 0);
  }
  j.set("exact_group_verifications", Json::u64(verifications));
  j.set("sketch_seconds", Json::f64(0.0));
  j.set("plan_seconds", Json::f64(0.0));
  j.set("build_seconds", Json::f64(build_seconds));
  j.set("input_document_count", Json::u64(m.colors.size()));
  j.set("memory_guard_split_groups", Json::u64(0));
  Json colors = Json::array();
  for (auto &c : m.colors) {
    Json o = Json::object();
    o.set("name", Json::str(c.name));
    o.set("path", Json::str(c.path));
    o.set("occurrences", Json::u64(c.occurrences));
    o.set("unique_kmers", Json::u64(c.unique_kmers));
    colors.push(std::move(o));
  }
  j.set("colors", std::move(colors));
  Json banks = Json::array();
  for (auto &bm : m.banks) {
    Json o = Json::object();
    o.set("file", Json::str(bm.file));
    o.set("slots", Json::u64(bm.slots));
    o.set("row_bits", Json::u64(bm.row_bits));
    Json gs = Json::array();
    for (auto &g : bm.groups) {
      Json go = Json::object();
      Json mem = Json::array();
      for (uint32_t x : g.members) mem.push(Json::u64(x));
      go.set("members", std::move(mem));
      go.set("width", Json::u64(g.width));
      go.set("offset", Json::u64(g.offset));
      go.set("union_cardinality", Json::u64(g.union_cardinality));
      go.set("fallback_keys", Json::u64(g.fallback_keys));
      go.set("abandoned_keys", Json::u64(g.abandoned_keys));
      gs.push(std::move(go));
    }
    o.set("groups", std::move(gs));
    o.set("fingerprint_first", Json::boolean(true));
    if (bm.routed) {
      Json r = Json::object();
      r.set("source_slots", Json::u64(bm.routing_source_slots));
      r.set("partition", Json::u64(bm.partition));
      r.set("partitions", Json::u64(bm.partitions));
      o.set("routing", std::move(r));
    }
    banks.push(std::move(o));
  }
  j.set("banks", std::move(banks));
  m.json = std::move(j);
  validate_manifest(m);
  idx->payload_loaded = true;
  idx->rec = std::move(rec);
  idx->have_rec = !cfg.drop_abandoned;
  (void)abandoned_total;
}
}  // namespace zg

extern "C" int zg_builder_finish(zg_builder *b, zg_index **out) {
  try {
    ensure(b && out, "null argument", -1);
    b->drain();
    ensure(!b->colors.empty(), "no datasets were added", -1);
    if (b->cfg.merge_group_size > 1) {
      // merged groups: consecutive datasets, built in parallel
      uint32_t gsz = b->cfg.merge_group_size;
      size_t n = b->colors.size();
      size_t n_groups = (n + gsz - 1) / gsz;
      b->groups.assign(n_groups, GroupResult());
      for (size_t gi = 0; gi < n_groups; gi++) {
        b->submit([b, gi, gsz, n] {
          std::vector<uint32_t> members;
          std::vector<std::vector<uint64_t> *> sets;
          for (size_t c = gi * gsz; c < std::min(n, (gi + 1) * gsz); c++) {
            members.push_back((uint32_t)c);
            sets.push_back(&b->pending[c]);
          }
          build_group(b->cfg, members, sets, b->groups[gi]);
        });
      }
      b->drain();
    }
    partition_oversized(b);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - b->start).count();
    auto idx = std::make_unique<zg_index>();
    finish_index(b, idx.get(), secs);
    delete b;
    *out = idx.release();
    return 0;
  } catch (...) {
    return translate_exception();
  }
}

extern "C" int zg_builder_destroy(zg_builder *b) {
  delete b;
  return 0;
}

// read_fof (src/lib.rs:323-381)
static std::vector<std::pair<std::string, std::string>> read_fof(const std::string &path) {
  std::string text = read_file(path);
  std::vector<std::pair<std::string, std::string>> out;
  size_t i = 0, line_no = 0;
  auto trim = [](std::string s) {
    size_t a = 0, e = s.size();
    while (a < e && isspace((unsigned char)s[a])) a++;
    while (e > a && isspace((unsigned char)s[e - 1])) e--;
    return s.substr(a, e - a);
  };
  while (i <= text.size()) {
    size_t j = text.find('\n', i);
    if (j == std::string::npos) j = text.size();
    std::string line = trim(text.substr(i, j - i));
    line_no++;
    i = j + 1;
    if (line.empty() || line[0] == '#') {
      if (j >= text.size()) break;
      continue;
    }
    std::string name, p;
    size_t tab = line.find('\t');
    if (tab != std::string::npos) {
      name = trim(line.substr(0, tab));
      size_t tab2 = line.find('\t', tab + 1);
      p = trim(line.substr(tab + 1, tab2 == std::string::npos ? std::string::npos : tab2 - tab - 1));
    } else {
      size_t sp = line.find_first_of(" \t\r\n\f\v");
      if (sp != std::string::npos) {
        name = line.substr(0, sp);
        size_t b2 = line.find_first_not_of(" \t\r\n\f\v", sp);
        size_t e2 = line.find_first_of(" \t\r\n\f\v", b2);
        p = line.substr(b2, e2 == std::string::npos ? std::string::npos : e2 - b2);
      } else {
        p = line;
        size_t slash = p.find_last_of('/');
        name = slash == std::string::npos ? p : p.substr(slash + 1);
        for (const char *ext : {".zstd", ".zst", ".gzip", ".gz", ".xz", ".bz2", ".bz", ".bzip2",
                                ".fasta", ".fastq", ".fna", ".fa", ".fq", ".fas"}) {
          size_t el = strlen(ext);
          if (name.size() >= el && name.compare(name.size() - el, el, ext) == 0)
            name.resize(name.size() - el);
        }
        if (name.empty()) name = "color";
      }
    }
    ensure(!name.empty(), (path + ":" + std::to_string(line_no) + " has an empty color name").c_str(), -1);
    out.push_back({name, p});
    if (j >= text.size()) break;
  }
  return out;
}

// fuse_many_indexes (src/power_merge/fusion.rs:233-480): concatenates
// compatible indexes without rereading their datasets. Bank files are copied
// under component-c- names (components numbered from 1 in input order),
// dataset ids of later inputs are offset, recovery maps are merged by hash
// (equal hashes keep input order, so origins stay sorted), and the manifest
// becomes version 10.
extern "C" int zg_fusion_command(const char *const *inputs, uint32_t n_inputs,
                                 const char *out_dir, int err_fd) {
  std::vector<zg_index *> opened;
  auto close_all = [&] {
    for (zg_index *x : opened) zg_index_close(x);
    opened.clear();
  };
  try {
    ensure(inputs && out_dir, "null argument", -1);
    ensure(n_inputs >= 2, "fusion requires at least two input indexes", -1);
    struct stat st;
    if (stat(out_dir, &st) == 0)
      fail(-1, std::string("output ") + out_dir + " already exists, choose a new directory");
    auto start = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < n_inputs; i++) {
      ensure(inputs[i] != nullptr, "null input", -1);
      zg_index *x = nullptr;
      int rc = zg_index_open(inputs[i], &x);
      if (rc) throw Error(rc, std::string("reading input index ") + inputs[i] + ": " + zg_last_error());
      opened.push_back(x);
    }
    const Manifest &m0 = opened[0]->m;
    for (zg_index *x : opened) {
      const Manifest &m = x->m;
      ensure(m.version >= 2, "fusion requires separate recovery files", -1);
      auto same = [&](bool eq, const char *field) {
        if (!eq) fail(-1, std::string("incompatible indexes: ") + field + " differs");
      };
      same(m.backend == m0.backend, "backend");
      same(m.k == m0.k, "k");
      same(m.index_mode == m0.index_mode, "index_mode");
      same(m.minimizer_size == m0.minimizer_size, "minimizer_size");
      same(m.minimizer_scheme == m0.minimizer_scheme, "minimizer_scheme");
      same(m.homopolymer_compression == m0.homopolymer_compression, "homopolymer_compression");
      for (const char *f : {"seed_pattern", "randstrobemers", "randstrobe_scheme"}) {
        std::string a, b;
        if (const Json *v = m.json.get(f)) json_write(a, *v);
        if (const Json *v = m0.json.get(f)) json_write(b, *v);
        same(a == b, f);
      }
      same(m.modimizer_sampling == m0.modimizer_sampling, "modimizer_sampling");
      same(m.findere_z == m0.findere_z, "findere_z");
      same(m.hashes == m0.hashes, "hashes");
      same(m.seed == m0.seed, "seed");
      same(m.fingerprint_bits == m0.fingerprint_bits, "fingerprint_bits");
      same(m.has_segment_length == m0.has_segment_length &&
               (!m.has_segment_length || m.segment_length == m0.segment_length),
           "segment_length");
      same(m.abandoned_recovery == m0.abandoned_recovery, "abandoned_recovery");
    }
    std::string stage = std::string(out_dir) + ".staging";
    ensure(mkdir(stage.c_str(), 0755) == 0, "cannot create staging directory", -2);
    Json out = m0.json;
    Json banks = Json::array(), colors = Json::array(), source_ids = Json::array(),
         skipped = Json::array();
    uint64_t offset = 0, source_offset = 0;
    uint64_t component = 0;
    std::vector<uint64_t> offsets;
    for (zg_index *x : opened) {
      const Manifest &m = x->m;
      offsets.push_back(offset);
      int64_t previous = -1;
      const Json *jb = m.json.get("banks");
      for (size_t bi = 0; bi < m.banks.size(); bi++) {
        BankMeta meta = m.banks[bi];
        if ((int64_t)meta.component != previous) {
          component++;
          previous = (int64_t)meta.component;
        }
        meta.component = component;
        const std::string name = meta.expected_filename();
        write_file(path_join(stage, name), read_file(path_join(x->dir, m.banks[bi].file)));
        Json b = jb->items[bi];
        Json nb = Json::object();
        nb.set("component", Json::u64(component));
        for (auto &f : b.fields)
          if (f.first != "component") nb.set(f.first, f.second);
        nb.set("file", Json::str(name));
        Json &groups = nb.set("groups", nb.get("groups") ? *nb.get("groups") : Json::array());
        for (Json &g : groups.items) {
          Json mem = Json::array();
          for (const Json &id : g.get("members")->items)
            mem.push(Json::u64(std::stoull(id.text) + offset));
          g.set("members", std::move(mem));
        }
        banks.push(std::move(nb));
      }
      if (const Json *c = m.json.get("colors"))
        for (const Json &cj : c->items) colors.push(cj);
      const Json *ids = m.json.get("source_document_ids");
      if (ids && ids->kind == Json::Array && !ids->items.empty()) {
        for (const Json &id : ids->items) source_ids.push(Json::u64(std::stoull(id.text) + source_offset));
      } else {
        for (size_t c = 0; c < m.colors.size(); c++) source_ids.push(Json::u64(c + source_offset));
      }
      if (const Json *sk = m.json.get("skipped_documents"))
        for (const Json &e : sk->items) {
          Json ne = e;
          if (const Json *sid = e.get("source_document_id"))
            ne.set("source_document_id", Json::u64(std::stoull(sid->text) + source_offset));
          skipped.push(std::move(ne));
        }
      uint64_t docs = m.colors.size();
      if (const Json *idc = m.json.get("input_document_count"))
        docs = std::max<uint64_t>(docs, std::stoull(idc->text));
      offset += m.colors.size();
      source_offset += docs;
      ensure(offset <= UINT32_MAX, "dataset IDs must fit in 32 bits", -1);
    }
    // recovery maps merged by hash
    uint64_t rec_hashes = 0, rec_origins = 0;
    if (m0.abandoned_recovery) {
      std::vector<RecoveryMap> maps;
      for (zg_index *x : opened) maps.push_back(read_recovery_map(x->dir, x->m));
      RecoveryMap merged;
      std::vector<size_t> at(maps.size(), 0);
      while (true) {
        uint64_t best = UINT64_MAX;
        for (size_t i = 0; i < maps.size(); i++)
          if (at[i] < maps[i].size()) best = std::min<uint64_t>(best, maps[i].hashes[at[i]]);
        if (best == UINT64_MAX) break;
        for (size_t i = 0; i < maps.size(); i++)
          if (at[i] < maps[i].size() && maps[i].hashes[at[i]] == best) {
            for (uint64_t j = maps[i].off[at[i]]; j < maps[i].off[at[i] + 1]; j++)
              merged.origins.push_back((uint32_t)(maps[i].origins[j] + offsets[i]));
            at[i]++;
          }
        merged.hashes.push_back((uint32_t)best);
        merged.off.push_back(merged.origins.size());
      }
      write_recovery_map(path_join(stage, kAbandonedFile), merged);
      rec_hashes = merged.size();
      rec_origins = merged.origins.size();
    }
    // manifest (fusion.rs:384-447)
    for (size_t i = 1; i < opened.size(); i++) {
      const Json &r = opened[i]->m.json;
      for (const char *f : {"planning_pair_evaluations", "planning_possible_pair_evaluations",
                            "promoted_groups", "singleton_logical_bits", "predicted_logical_bits",
                            "planned_groups", "rejected_groups", "overflow_groups",
                            "exact_group_verifications", "memory_guard_split_groups"}) {
        const Json *a = out.get(f), *b = r.get(f);
        if (a && b && a->kind == Json::Number && b->kind == Json::Number)
          out.set(f, Json::u64(std::stoull(a->text) + std::stoull(b->text)));
      }
      for (const char *f : {"sketch_seconds", "plan_seconds"}) {
        const Json *a = out.get(f), *b = r.get(f);
        if (a && b && a->kind == Json::Number && b->kind == Json::Number)
          out.set(f, Json::f64(std::stod(a->text) + std::stod(b->text)));
      }
      const Json *a = out.get("max_subindex_size"), *b = r.get("max_subindex_size");
      if (a && b && a->kind == Json::Number && b->kind == Json::Number)
        out.set("max_subindex_size",
                Json::u64(std::max(std::stoull(a->text), std::stoull(b->text))));
      else
        out.set("max_subindex_size", Json::null());
    }
    out.set("version", Json::u64(10));
    out.set("abandoned_hashes", Json::u64(rec_hashes));
    out.set("abandoned_origins", Json::u64(rec_origins));
    out.set("colors", std::move(colors));
    out.set("banks", std::move(banks));
    out.set("source_document_ids", std::move(source_ids));
    out.set("skipped_documents", std::move(skipped));
    out.set("input_document_count", Json::u64(source_offset));
    out.set("max_subindexes", Json::null());
    out.set("min_subindex_documents", Json::u64(1));
    out.set("merge_memory_budget_bytes", Json::null());
    out.set("construction_cell_storage", Json::str("copied-compressed-banks"));
    out.set("column_order_method", Json::null());
    out.set("construction_abandoned_storage",
            Json::str(m0.abandoned_recovery ? "streaming-merge" : "omitted"));
    out.set("build_seconds",
            Json::f64(std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count()));
    std::string text;
    json_write(text, out);
    write_file(path_join(stage, "manifest.json"), text);
    close_all();
    ensure(rename(stage.c_str(), out_dir) == 0, "cannot publish the index directory", -2);
    // the fused index must pass the full validation
    zg_index *check = nullptr;
    int rc = zg_index_open(out_dir, &check);
    if (check) zg_index_close(check);
    if (rc) return rc;
    if (err_fd >= 0) {
      std::string msg = "fused " + std::to_string(n_inputs) + " indexes: " + std::to_string(offset) +
                        " datasets, " + std::to_string(component) + " components\n";
      (void)!write(err_fd, msg.data(), msg.size());
    }
    return 0;
  } catch (...) {
    close_all();
    return translate_exception();
  }
}

extern "C" int zg_build_command(const zg_build_config *cfg, const char *fof, const char *out_dir,
                                int err_fd) {
  zg_builder *b = nullptr;
  zg_index *idx = nullptr;
  try {
    ensure(cfg && fof && out_dir, "null argument", -1);
    struct stat st;
    ensure(stat(out_dir, &st) != 0, "output directory already exists", -1);
    auto inputs = read_fof(fof);
    ensure(!inputs.empty(), "the file-of-files lists no datasets", -1);
    int rc = zg_builder_create(cfg, &b);
    if (rc) return rc;
    for (auto &in : inputs) {
      Records r = parse_sequences(read_maybe_compressed(in.second));
      rc = zg_builder_add_color(b, in.first.c_str(), (const uint8_t *)r.seq.data(), r.off.data(),
                                r.size());
      if (rc) {
        zg_builder_destroy(b);
        return rc;
      }
    }
    rc = zg_builder_finish(b, &idx);
    b = nullptr;
    if (rc) return rc;
    // keep the FOF paths in the manifest
    Json *colors = nullptr;
    for (auto &f : idx->m.json.fields)
      if (f.first == "colors") colors = &f.second;
    for (size_t i = 0; colors && i < colors->items.size(); i++)
      colors->items[i].set("path", Json::str(inputs[i].second));
    rc = zg_index_write(idx, out_dir);
    if (rc == 0 && err_fd >= 0) {
      size_t groups = 0;
      for (auto &bk : idx->m.banks) groups += bk.groups.size();
      std::string line = "built " + std::string(out_dir) + ": banks=" +
                         std::to_string(idx->m.banks.size()) + ", groups=" + std::to_string(groups) +
                         ", abandoned_hashes=" + std::to_string(idx->m.abandoned_hashes) + "\n";
      ssize_t w = write(err_fd, line.data(), line.size());
      (void)w;
    }
    zg_index_close(idx);
    return rc;
  } catch (...) {
    if (b) zg_builder_destroy(b);
    if (idx) zg_index_close(idx);
    return translate_exception();
  }
}

extern "C" int zg_index_write(const zg_index *idx, const char *out_dir) {
  try {
    ensure(idx && out_dir, "null argument", -1);
    ensure(idx->m.version != 1, "writing version-1 indexes is not supported", -6);
    zg_index *ix = const_cast<zg_index *>(idx);  // host copies are decoded on first use
    if (ix->m.abandoned_recovery && !ix->have_rec) {
      ensure(!ix->dir.empty(), "recovery map not in host memory", -1);
      ix->rec = read_recovery_map(ix->dir, ix->m);
      ix->have_rec = true;
    }
    std::string stage = std::string(out_dir) + ".staging";
    ensure(mkdir(stage.c_str(), 0755) == 0, "cannot create staging directory", -2);
    for (size_t i = 0; i < idx->m.banks.size(); i++) {
      const std::vector<uint8_t> &p = host_payload(ix, i);
      write_bank_file(path_join(stage, idx->m.banks[i].file), idx->m.banks[i], p.data(), p.size(),
                      (int)zg_host_threads());
    }
    if (idx->m.abandoned_recovery)
      write_recovery_map(path_join(stage, kAbandonedFile), idx->rec);
    std::string json;
    json_write(json, idx->m.json);
    write_file(path_join(stage, "manifest.json"), json);
    ensure(rename(stage.c_str(), out_dir) == 0, "cannot publish the index directory", -2);
    return 0;
  } catch (...) {
    return translate_exception();
  }
}
