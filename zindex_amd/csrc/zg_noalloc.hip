// zg_noalloc.hip -- zg_query_reads without device allocation.
//
// Callers that may not allocate device memory (tight serving loops,
// harnesses that forbid hipMalloc) run the read-ordered lookup out of one static device array per
// GPU. Every working buffer of the query -- the reads and their offsets, the
// packed bases and invalid bits, minimizer flags, keys and scan levels, the
// per-read totals and counts, per-wave partial sums, the aggregate and the
// work counters -- is carved out of it; the caller's host buffers are the
// only other memory touched. The lookup kernels are the same launches
// zg_query_reads makes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "zg_internal.h"

namespace zg {
namespace {

// 64 MiB, held by every GPU the library touches until the process ends.
constexpr uint64_t kScratchBytes = 64ULL << 20;
// Per-wave partial sums (aggregate-only queries) take at most this much; the
// grid is narrowed to fit.
constexpr uint64_t kPartialBytes = 16ULL << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_setup;
uint8_t *g_scratch[kMaxDevices];  // device address of `scratch` on each GPU

// Address of this GPU's copy of the array, looked up once per device with
// that device current (the array's host-side name is not a device address).
uint8_t *scratch_on(int dev) {
  ensure(dev >= 0 && dev < kMaxDevices, "device ordinal out of range", ZG_ERR_DEVICE);
  std::lock_guard<std::mutex> lock(g_setup);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    ensure(p != nullptr, "hipGetSymbolAddress returned no address", ZG_ERR_DEVICE);
    g_scratch[dev] = (uint8_t *)p;
  }
  return g_scratch[dev];
}

// Makes dev current for a scope and restores the caller's device after.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    check_hip(hipGetDevice(&prev), "hipGetDevice");
    if (prev != dev) check_hip(hipSetDevice(dev), "hipSetDevice");
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Bump carving of the scratch array in 256-byte steps.
struct Carver {
  uint8_t *base;
  uint64_t used = 0;
  explicit Carver(uint8_t *b) : base(b) {}
  static uint64_t round(uint64_t n) { return (std::max<uint64_t>(n, 1) + 255) / 256 * 256; }
  void take(DevBuf &d, uint64_t n) {
    n = round(n);
    ensure(used + n <= kScratchBytes,
           "query exceeds the 64 MiB scratch of zg_query_reads_noalloc (split the reads)",
           ZG_ERR_INVALID);
    d.borrow(base + used, n);
    used += n;
  }
};

// roff[r] = raw[r] - r0, poff[r] = roff[r] + r (one separator per read).
__global__ void rebase_offsets_kernel(uint64_t *roff, uint64_t *poff, uint64_t n, uint64_t r0) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = roff[i] - r0;
    roff[i] = v;
    poff[i] = v + i;
  }
}

// Bytes of the buffers sized before seeding (the minimizer keys come after
// their count is known).
uint64_t planned_bytes(const Manifest &m, uint64_t n_reads, uint64_t bases, uint64_t positions,
                       bool records, uint64_t waves, size_t banks) {
  const uint64_t C = m.colors.size();
  uint64_t t = Carver::round(bases + 64) + 3 * Carver::round((n_reads + 1) * 8) +
               Carver::round(code_words(positions, m.k) * 8) +
               Carver::round(invalid_words(positions, m.k) * 8) + Carver::round(C * 8 + 8) +
               Carver::round((banks + 1) * 4);
  t += records ? Carver::round(n_reads * C * 4) : Carver::round(waves * C * 8);
  if (m.index_mode == 1) {
    const uint64_t blocks = (positions + ocm_block_positions() - 1) / ocm_block_positions();
    t += 2 * Carver::round(positions + 64) + Carver::round((blocks + 1) * 4) +
         Carver::round((blocks + 1) * 8) + Carver::round(scan_temp_bytes_u32_to_u64(blocks + 1)) +
         Carver::round((n_reads + 1) * 8);
  }
  return t;
}

}  // namespace
}  // namespace zg

using namespace zg;

extern "C" int zg_query_reads_noalloc(zg_index *idx, const uint8_t *seq, const uint64_t *read_off,
                                      uint64_t n_reads, uint32_t flags, uint64_t *totals,
                                      uint32_t *counts, uint64_t *aggregate) {
  try {
    ensure(idx && (n_reads == 0 || read_off), "null argument", ZG_ERR_INVALID);
    ensure(!counts || (flags & ZG_QUERY_RECORDS), "per-read counts need ZG_QUERY_RECORDS",
           ZG_ERR_INVALID);
    ensure(idx->on_device, "index is not resident on the device (call zg_index_load)",
           ZG_ERR_INVALID);
    const Manifest &m = idx->m;
    ensure(!m.has_randstrobemers && m.index_mode != 3 && !m.homopolymer_compression &&
               (m.index_mode == 1 || m.k <= 32),
           "zg_query_reads_noalloc covers k-mers (k <= 32), minimizers and modimizers without "
           "homopolymer compression; use zg_query_reads",
           ZG_ERR_UNSUPPORTED);
    ensure(!any_streamed(idx), "zg_query_reads_noalloc needs every subindex resident in HBM",
           ZG_ERR_UNSUPPORTED);
    const uint32_t C = (uint32_t)m.colors.size();
    const uint64_t r0 = n_reads ? read_off[0] : 0;
    for (uint64_t r = 1; r <= n_reads; r++)
      ensure(read_off[r] >= read_off[r - 1], "read offsets must be non-decreasing",
             ZG_ERR_INVALID);
    const uint64_t bases = n_reads ? read_off[n_reads] - r0 : 0;
    ensure(bases == 0 || seq, "null sequence", ZG_ERR_INVALID);
    if (n_reads == 0) {
      if (aggregate) memset(aggregate, 0, C * 8ULL);
      return 0;
    }
    const uint64_t positions = bases + n_reads;

    DeviceScope on(idx->device);
    uint8_t *base = scratch_on(idx->device);
    hipStream_t s = idx->stream;

    zg_batch b;  // borrows every buffer; owns nothing on the device
    b.idx = idx;
    b.stream = s;
    b.flags = flags;
    b.records = (flags & ZG_QUERY_RECORDS) != 0;
    b.n_reads = n_reads;
    b.bases = bases;
    b.positions = positions;
    // launch shape (the grid narrowed so per-wave partial sums fit their
    // share); configured again once the seed count is known
    auto shape = [&]() {
      configure(idx, &b);
      if (!b.records) {
        const uint64_t fit = std::max<uint64_t>(1, kPartialBytes / (4ULL * C * 8));
        b.kcfg.blocks = (uint32_t)std::min<uint64_t>(b.kcfg.blocks, fit);
      }
    };
    shape();
    const uint64_t waves = b.records ? 0 : (uint64_t)b.kcfg.blocks * 4;
    ensure(planned_bytes(m, n_reads, bases, positions, b.records, waves, idx->dbanks.size()) <=
               kScratchBytes,
           "query exceeds the 64 MiB scratch of zg_query_reads_noalloc (split the reads)",
           ZG_ERR_INVALID);

    Carver cv(base);
    cv.take(b.seq, bases + 64);
    cv.take(b.read_off, (n_reads + 1) * 8);
    cv.take(b.pos_off, (n_reads + 1) * 8);
    cv.take(b.codes, code_words(positions, m.k) * 8);
    cv.take(b.invalid, invalid_words(positions, m.k) * 8);
    cv.take(b.totals, (n_reads + 1) * 8);
    cv.take(b.agg, C * 8ULL + 8);
    cv.take(b.queue, (idx->dbanks.size() + 1) * 4);
    if (b.records) cv.take(b.counts, n_reads * (uint64_t)C * 4);
    else cv.take(b.partial, waves * C * 8);
    b.waves = waves;

    // reads in, offsets rebased to the batch
    check_hip(hipMemcpyAsync(b.seq.p, seq + r0, bases, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    check_hip(hipMemcpyAsync(b.read_off.p, read_off, (n_reads + 1) * 8, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync");
    const uint64_t grid = std::min<uint64_t>((n_reads + 1 + 255) / 256, 4096);
    rebase_offsets_kernel<<<(uint32_t)grid, 256, 0, s>>>(b.read_off.as<uint64_t>(),
                                                         b.pos_off.as<uint64_t>(), n_reads + 1, r0);
    check_hip(hipGetLastError(), "rebase_offsets_kernel");

    // seeds (the steps of seed_batch for these modes)
    launch_fill_u64(b.codes.as<uint64_t>(), code_words(positions, m.k), 0, s);
    launch_fill_u64(b.invalid.as<uint64_t>(), invalid_words(positions, m.k), ~0ULL, s);
    launch_pack(b.seq.as<uint8_t>(), b.read_off.as<uint64_t>(), b.pos_off.as<uint64_t>(), n_reads,
                positions, b.codes.as<uint64_t>(), b.invalid.as<uint64_t>(), s);
    launch_fill_u64(b.totals.as<uint64_t>(), n_reads, 0, s);
    if (m.index_mode != 1) {
      launch_count_kmers(b.invalid.as<uint64_t>(), b.codes.as<uint64_t>(), m.seed,
                         m.index_mode == 2 ? m.modimizer_sampling : 0, b.pos_off.as<uint64_t>(),
                         n_reads, positions, (uint32_t)m.k, b.totals.as<uint64_t>(), s);
    } else {
      const uint64_t blocks = (positions + ocm_block_positions() - 1) / ocm_block_positions();
      cv.take(b.flag, positions + 64);
      cv.take(b.delta, positions + 64);
      cv.take(b.block_counts, (blocks + 1) * 4);
      cv.take(b.block_off, (blocks + 1) * 8);
      const size_t tb = scan_temp_bytes_u32_to_u64(blocks + 1);
      cv.take(b.scan_temp, tb);
      cv.take(b.key_off, (n_reads + 1) * 8);
      launch_ocm_select(b.codes.as<uint64_t>(), b.invalid.as<uint64_t>(), b.pos_off.as<uint64_t>(),
                        n_reads, positions, (uint32_t)m.k, (uint32_t)m.minimizer_size, m.seed,
                        idx->inner_table.as<uint64_t>(), b.flag.as<uint8_t>(),
                        b.delta.as<uint8_t>(), b.block_counts.as<uint32_t>(), nullptr, s);
      launch_fill_u32(b.block_counts.as<uint32_t>() + blocks, 1, 0, s);
      launch_exclusive_scan_u32_to_u64(b.block_counts.as<uint32_t>(), b.block_off.as<uint64_t>(),
                                       blocks + 1, b.scan_temp.p, tb, s);
      uint64_t total = 0;
      check_hip(hipMemcpyAsync(&total, b.block_off.as<uint64_t>() + blocks, 8,
                               hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
      check_hip(hipStreamSynchronize(s), "seed");
      cv.take(b.keys, (total + 64) * 8);
      cv.take(b.key_win, (total + 64) * 8);
      launch_ocm_compact(b.codes.as<uint64_t>(), b.flag.as<uint8_t>(), b.delta.as<uint8_t>(),
                         b.block_off.as<uint64_t>(), positions, (uint32_t)m.minimizer_size,
                         idx->spaced, b.keys.as<uint64_t>(), b.key_win.as<uint64_t>(), s);
      launch_segment_bounds(b.key_win.as<uint64_t>(), total, b.pos_off.as<uint64_t>(), n_reads,
                            b.key_off.as<uint64_t>(), s);
      launch_diff(b.key_off.as<uint64_t>(), n_reads, b.totals.as<uint64_t>(), s);
      b.n_keys = total;
      b.explicit_keys = true;
    }

    shape();  // span for the seed count; the grid is unchanged
    ensure(b.records || (uint64_t)b.kcfg.blocks * 4 == waves, "launch shape changed",
           ZG_ERR_INVALID);

    // outputs zeroed, one persistent lookup per bank, per-colour sums
    if (b.records) launch_fill_u32(b.counts.as<uint32_t>(), n_reads * (uint64_t)C, 0, s);
    else launch_fill_u64(b.partial.as<uint64_t>(), waves * C, 0, s);
    launch_fill_u64(b.agg.as<uint64_t>(), C, 0, s);
    launch_fill_u32(b.queue.as<uint32_t>(), idx->dbanks.size(), 0, s);
    if (b.explicit_keys ? b.n_keys > 0 : positions > 0)
      for (size_t i = 0; i < idx->dbanks.size(); i++) launch_lookup(lookup_args(idx, &b, i), b.kcfg, s);
    if (b.records) launch_colsum(b.counts.as<uint32_t>(), n_reads, C, b.agg.as<uint64_t>(), s);
    else launch_reduce_partials(b.partial.as<uint64_t>(), waves, C, b.agg.as<uint64_t>(), s);
    check_hip(hipGetLastError(), "lookup");

    if (totals)
      check_hip(hipMemcpyAsync(totals, b.totals.p, n_reads * 8, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    if (counts)
      check_hip(hipMemcpyAsync(counts, b.counts.p, n_reads * (uint64_t)C * 4, hipMemcpyDeviceToHost,
                               s),
                "hipMemcpyAsync");
    if (aggregate)
      check_hip(hipMemcpyAsync(aggregate, b.agg.p, C * 8ULL, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "query");
    return 0;
  } catch (...) {
    return translate_exception();
  }
}
