// GPU check of the sorted-sweep building blocks against a CPU restatement:
// the radix sort of (hash high half, hash low half | position) entries, run
// heads, the scan and assign to distinct-seed ids, the sort back to position
// order, and the per-read hit-row sums (records and per-colour totals).
// Built by __graft_entry__.build() against libzindex_amd.so; run by
// tests/test_gpu_parity.py. Prints one line per step and exits 0 when every
// step matches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

#include "zg_kernels.h"

namespace {

template <class T>
T *dev(size_t n) {
  T *p = nullptr;
  zg::check_hip(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
  zg::check_hip(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)), "hipMemset");
  return p;
}
template <class T>
std::vector<T> get(const T *p, size_t n) {
  std::vector<T> v(n);
  zg::check_hip(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  return v;
}
template <class T>
void put(T *p, const std::vector<T> &v) {
  zg::check_hip(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // namespace

static int run() {
  std::mt19937_64 rng(7);
  const uint64_t n_pos = 120000, n = 100000;
  // first hashes drawn from a small pool, so runs of equal seeds occur
  std::vector<uint64_t> pool(30000);
  for (auto &x : pool) x = rng();
  std::vector<uint64_t> hash(n);
  for (auto &x : hash) x = pool[rng() % pool.size()];
  std::vector<uint32_t> all(n_pos);
  std::iota(all.begin(), all.end(), 0);
  std::shuffle(all.begin(), all.end(), rng);
  std::vector<uint32_t> pos(all.begin(), all.begin() + n);
  std::sort(pos.begin(), pos.end());  // listed in position order, as the list pass emits them
  std::vector<uint32_t> hi(n);
  std::vector<uint64_t> lp(n);
  for (uint64_t i = 0; i < n; i++) {
    hi[i] = (uint32_t)(hash[i] >> 32);
    lp[i] = (hash[i] << 32) | pos[i];
  }

  // 1. sort by the high half, stable
  uint32_t *dh = dev<uint32_t>(n), *dh2 = dev<uint32_t>(n);
  uint64_t *dp = dev<uint64_t>(n), *dp2 = dev<uint64_t>(n);
  put(dh, hi);
  put(dp, lp);
  const size_t st = std::max(zg::sort_pairs_temp_bytes(n), zg::sort_pairs_u32_temp_bytes(n));
  void *temp = dev<uint8_t>(st);
  zg::launch_sort_pairs(temp, st, dh, dh2, dp, dp2, n, nullptr);
  zg::check_hip(hipDeviceSynchronize(), "sort");
  std::vector<uint32_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return hi[a] < hi[b]; });
  const auto h2 = get(dh2, n);
  const auto p2 = get(dp2, n);
  size_t bad_sort = 0;
  for (uint64_t i = 0; i < n; i++) bad_sort += h2[i] != hi[ord[i]] || p2[i] != lp[ord[i]];
  printf("sort mismatches %zu\n", bad_sort);

  // 2. heads, scan, assign, sort back to position order
  uint32_t *flags = dev<uint32_t>(n + 1), *apos = dev<uint32_t>(n), *aid = dev<uint32_t>(n);
  uint32_t *spos = dev<uint32_t>(n), *sid = dev<uint32_t>(n);
  uint64_t *off = dev<uint64_t>(n + 1), *first = dev<uint64_t>(n);
  const size_t tb = zg::scan_temp_bytes_u32_to_u64(n + 1);
  void *stemp = dev<uint8_t>(tb);
  zg::launch_sweep_heads(dh2, dp2, n, flags, nullptr);
  zg::launch_exclusive_scan_u32_to_u64(flags, off, n + 1, stemp, tb, nullptr);
  zg::launch_sweep_assign(flags, off, dh2, dp2, n, first, apos, aid, nullptr);
  int pos_bits = 1;
  while (pos_bits < 32 && (1ULL << pos_bits) < n_pos) pos_bits++;
  zg::launch_sort_pairs_u32(temp, st, apos, spos, aid, sid, n, pos_bits, nullptr);
  zg::check_hip(hipDeviceSynchronize(), "assign");
  const auto hoff = get(off, n + 1);
  const auto hfirst = get(first, n);
  const auto hspos = get(spos, n);
  const auto hsid = get(sid, n);
  const uint64_t distinct = hoff[n];
  // CPU: runs of equal hashes over the stably sorted list
  std::vector<uint32_t> eslot(n_pos, zg::kNone);
  std::vector<uint64_t> efirst;
  for (uint64_t i = 0; i < n; i++) {
    const uint64_t h = hash[ord[i]];
    if (i == 0 || h != hash[ord[i - 1]]) efirst.push_back(h);
    eslot[pos[ord[i]]] = (uint32_t)(efirst.size() - 1);
  }
  size_t bad_ids = distinct != efirst.size();
  for (uint64_t u = 0; u < std::min<uint64_t>(distinct, efirst.size()); u++)
    bad_ids += hfirst[u] != efirst[u];
  uint64_t e = 0;
  for (uint64_t p = 0; p < n_pos; p++)
    if (eslot[p] != zg::kNone) {
      bad_ids += e >= n || hspos[e] != p || hsid[e] != eslot[p];
      e++;
    }
  bad_ids += e != n;
  printf("distinct %llu (cpu %zu), id mismatches %zu\n", (unsigned long long)distinct,
         efirst.size(), bad_ids);

  // 3. per-read sums of random hit rows (CPU ids), records and totals
  const uint32_t region_bits = 1000, row_words = 40, region_off = 3, C = 1100;
  std::vector<uint32_t> rows(efirst.size() * row_words);
  for (auto &x : rows) x = (uint32_t)rng() & (uint32_t)rng();
  std::vector<uint32_t> bit_color(region_bits);
  for (uint32_t j = 0; j < region_bits; j++) bit_color[j] = j % 7 == 3 ? zg::kNone : (j * 13) % C;
  std::vector<uint64_t> read_off{0};  // reads over [0, n_pos), some empty
  while (read_off.back() < n_pos)
    read_off.push_back(std::min<uint64_t>(n_pos, read_off.back() + (rng() % 3 ? rng() % 3000 : 0)));
  const uint64_t n_reads = read_off.size() - 1;
  std::vector<uint32_t> epos, eid;
  for (uint64_t p = 0; p < n_pos; p++)
    if (eslot[p] != zg::kNone) {
      epos.push_back((uint32_t)p);
      eid.push_back(eslot[p]);
    }
  uint32_t *drows = dev<uint32_t>(rows.size()), *dbc = dev<uint32_t>(region_bits);
  uint32_t *counts = dev<uint32_t>(n_reads * C), *dpos = dev<uint32_t>(epos.size());
  uint32_t *did = dev<uint32_t>(eid.size());
  uint64_t *doff = dev<uint64_t>(read_off.size()), *partial = dev<uint64_t>(C);
  put(drows, rows);
  put(dbc, bit_color);
  put(doff, read_off);
  put(dpos, epos);
  put(did, eid);
  zg::launch_sweep_sum(drows, row_words, region_off, region_bits, dbc, dpos, did, epos.size(),
                       doff, n_reads, 0, 1, counts, partial, C, nullptr);
  zg::launch_sweep_sum(drows, row_words, region_off, region_bits, dbc, dpos, did, epos.size(),
                       doff, n_reads, 0, 0, counts, partial, C, nullptr);
  zg::check_hip(hipDeviceSynchronize(), "sum");
  const auto hc = get(counts, n_reads * C);
  const auto hp = get(partial, C);
  std::vector<uint32_t> ec(n_reads * C, 0);
  std::vector<uint64_t> ep(C, 0);
  for (uint64_t r = 0; r < n_reads; r++)
    for (uint64_t p = read_off[r]; p < read_off[r + 1]; p++) {
      const uint32_t u = eslot[p];
      if (u == zg::kNone) continue;
      for (uint32_t j = 0; j < region_bits; j++)
        if ((rows[u * row_words + region_off + j / 32] >> (j % 32)) & 1 && bit_color[j] != zg::kNone) {
          ec[r * C + bit_color[j]]++;
          ep[bit_color[j]]++;
        }
    }
  size_t bad_counts = 0, bad_totals = 0;
  for (size_t i = 0; i < ec.size(); i++) bad_counts += ec[i] != hc[i];
  for (uint32_t c = 0; c < C; c++) bad_totals += ep[c] != hp[c];
  printf("sum: reads %llu, count mismatches %zu, total mismatches %zu\n",
         (unsigned long long)n_reads, bad_counts, bad_totals);
  const bool ok = !bad_sort && !bad_ids && !bad_counts && !bad_totals && distinct > 0;
  printf("%s\n", ok ? "sweep pieces ok" : "sweep pieces FAILED");
  return ok ? 0 : 1;
}

int main() {
  try {
    return run();
  } catch (const std::exception &e) {
    printf("sweep pieces error: %s\n", e.what());
    return 2;
  }
}
