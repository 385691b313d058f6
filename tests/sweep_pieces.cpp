// GPU check of the sorted sweep's building blocks against a CPU restatement:
// the exclusive scans (one, two and three levels of tiles), the stable radix
// passes that order entries by their first segment mulhi(h, count) (one,
// two and three passes), the segment starts, and the per-read hit-row sums
// (records and per-colour totals, ragged and empty reads, rows narrower and
// wider than a wave). Built by __graft_entry__.build() against
// libzindex_amd.so; run by tests/test_gpu_parity.py. Prints one line per
// step and exits 0 when every step matches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <exception>
#include <numeric>
#include <random>
#include <vector>

#include "zg_kernels.h"

namespace zg {
int sweep_sort_entries(const uint64_t *h, const uint32_t *p, const uint64_t *n_dev, uint64_t n_max,
                       uint64_t count, uint64_t *h1, uint32_t *p1, uint64_t *h2, uint32_t *p2,
                       hipStream_t s);
void sweep_bounds(const uint64_t *h, const uint64_t *n_dev, uint64_t n_max, uint64_t count,
                  uint32_t *start, hipStream_t s);
void sweep_sum_rows(uint32_t *hits, uint32_t row_words, const uint32_t *word_color,
                    const uint64_t *listed, const uint64_t *pos_off, uint64_t n_reads, uint64_t p0,
                    uint64_t n_pos, uint32_t records, uint32_t *counts, uint64_t *partial,
                    uint32_t n_colors, hipStream_t s);
}  // namespace zg

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
uint64_t seg_of(uint64_t h, uint64_t count) {
  return (uint64_t)(((unsigned __int128)h * count) >> 64);
}

int check_scan(std::mt19937_64 &rng) {
  int bad = 0;
  for (uint64_t n : {1ULL, 4095ULL, 4096ULL, 4097ULL, (1ULL << 20) + 3, (1ULL << 24) + 5}) {
    std::vector<uint32_t> in(n);
    for (auto &x : in) x = (uint32_t)(rng() % 1000);
    std::vector<uint64_t> want(n);
    uint64_t run = 0;
    for (uint64_t i = 0; i < n; i++) {
      want[i] = run;
      run += in[i];
    }
    uint32_t *din = dev<uint32_t>(n);
    uint64_t *dout = dev<uint64_t>(n);
    put(din, in);
    const size_t tb = zg::scan_temp_bytes_u32_to_u64(n);
    void *temp = dev<uint8_t>(tb);
    zg::launch_exclusive_scan_u32_to_u64(din, dout, n, temp, tb, 0);
    auto got = get(dout, n);
    // the 64-bit form in place on the same values
    std::vector<uint64_t> in64(in.begin(), in.end());
    put(dout, in64);
    zg::launch_exclusive_scan_u64(dout, dout, n, temp, tb, 0);
    auto got64 = get(dout, n);
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++) m += (got[i] != want[i]) + (got64[i] != want[i]);
    printf("scan n=%llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)m);
    bad += m != 0;
    (void)hipFree(din);
    (void)hipFree(dout);
    (void)hipFree(temp);
  }
  return bad;
}

int check_sort_and_bounds(std::mt19937_64 &rng) {
  int bad = 0;
  const uint64_t n_max = 200000, n = 180000;
  for (uint64_t count : {3ULL, 131069ULL, (1ULL << 20) + 7}) {
    // first hashes from a small pool, so equal seeds and crowded segments occur
    std::vector<uint64_t> pool(40000);
    for (auto &x : pool) x = rng();
    std::vector<uint64_t> h(n_max);
    std::vector<uint32_t> p(n_max);
    for (uint64_t i = 0; i < n_max; i++) {
      h[i] = pool[rng() % pool.size()];
      p[i] = (uint32_t)(3 * i + 1);  // in position order, as the first pass lists them
    }
    std::vector<uint64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
      return seg_of(h[a], count) < seg_of(h[b], count);
    });
    uint64_t *dh = dev<uint64_t>(n_max), *h1 = dev<uint64_t>(n_max), *h2 = dev<uint64_t>(n_max);
    uint32_t *dp = dev<uint32_t>(n_max), *p1 = dev<uint32_t>(n_max), *p2 = dev<uint32_t>(n_max);
    uint64_t *dn = dev<uint64_t>(1);
    put(dh, h);
    put(dp, p);
    put(dn, std::vector<uint64_t>{n});
    const int which = zg::sweep_sort_entries(dh, dp, dn, n_max, count, h1, p1, h2, p2, 0);
    auto gh = get(which ? h2 : h1, n);
    auto gp = get(which ? p2 : p1, n);
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++) m += gh[i] != h[order[i]] || gp[i] != p[order[i]];
    printf("sort count=%llu n=%llu mismatches %llu\n", (unsigned long long)count,
           (unsigned long long)n, (unsigned long long)m);
    bad += m != 0;
    uint32_t *start = dev<uint32_t>(count + 1);
    zg::sweep_bounds(which ? h2 : h1, dn, n_max, count, start, 0);
    auto gs = get(start, count + 1);
    uint64_t mb = 0, at = 0;
    for (uint64_t s = 0; s <= count; s++) {
      while (at < n && seg_of(h[order[at]], count) < s) at++;
      mb += gs[s] != at;
    }
    printf("bounds count=%llu mismatches %llu\n", (unsigned long long)count,
           (unsigned long long)mb);
    bad += mb != 0;
    for (void *q : {(void *)dh, (void *)h1, (void *)h2, (void *)dp, (void *)p1, (void *)p2,
                    (void *)dn, (void *)start})
      (void)hipFree(q);
  }
  return bad;
}

int check_sum(std::mt19937_64 &rng) {
  int bad = 0;
  const uint32_t n_colors = 700;
  for (uint32_t row_words : {1u, 32u, 48u, 70u}) {
    // reads: ragged, some empty; the chunk covers reads [r0, r1)
    const uint64_t n_reads = 300, r0 = 37, r1 = 260;
    std::vector<uint64_t> pos_off(n_reads + 1, 0);
    for (uint64_t r = 0; r < n_reads; r++)
      pos_off[r + 1] = pos_off[r] + (r % 17 == 5 ? 1 : 1 + rng() % 9000);  // +1: separator
    const uint64_t p0 = pos_off[r0], n_pos = pos_off[r1] - p0;
    std::vector<uint32_t> hits(n_pos * row_words);
    for (auto &x : hits) x = (uint32_t)rng() & (uint32_t)rng();
    std::vector<uint64_t> listed((n_pos + 63) / 64 + 1);
    for (auto &x : listed) x = rng() | rng();
    std::vector<uint32_t> color(row_words * 32);
    for (uint32_t j = 0; j < row_words * 32; j++)
      color[j] = (j % 11 == 3) ? 0xFFFFFFFFu : (uint32_t)((j * 7) % n_colors);
    std::vector<uint64_t> want(n_reads * n_colors, 0), want_agg(n_colors, 0);
    uint64_t r = r0;
    for (uint64_t q = 0; q < n_pos; q++) {
      while (pos_off[r + 1] <= p0 + q) r++;
      if (!((listed[q >> 6] >> (q & 63)) & 1)) continue;
      for (uint32_t j = 0; j < row_words * 32; j++)
        if (((hits[q * row_words + j / 32] >> (j % 32)) & 1) && color[j] != 0xFFFFFFFFu) {
          want[r * n_colors + color[j]]++;
          want_agg[color[j]]++;
        }
    }
    uint32_t *dhits = dev<uint32_t>(hits.size()), *dcolor = dev<uint32_t>(color.size());
    uint64_t *dlisted = dev<uint64_t>(listed.size()), *doff = dev<uint64_t>(pos_off.size());
    uint32_t *dcounts = dev<uint32_t>(n_reads * n_colors);
    uint64_t *dpart = dev<uint64_t>(n_colors);
    put(dhits, hits);
    put(dcolor, color);
    put(dlisted, listed);
    put(doff, pos_off);
    zg::sweep_sum_rows(dhits, row_words, dcolor, dlisted, doff, n_reads, p0, n_pos, 1, dcounts,
                       nullptr, n_colors, 0);
    zg::sweep_sum_rows(dhits, row_words, dcolor, dlisted, doff, n_reads, p0, n_pos, 0, nullptr,
                       dpart, n_colors, 0);
    auto gc = get(dcounts, n_reads * n_colors);
    auto ga = get(dpart, n_colors);
    uint64_t mc = 0, ma = 0;
    for (uint64_t i = 0; i < n_reads * n_colors; i++) mc += gc[i] != want[i];
    for (uint32_t c = 0; c < n_colors; c++) ma += ga[c] != want_agg[c];
    printf("sum row_words=%u count mismatches %llu, total mismatches %llu\n", row_words,
           (unsigned long long)mc, (unsigned long long)ma);
    bad += mc != 0 || ma != 0;
    for (void *q : {(void *)dhits, (void *)dcolor, (void *)dlisted, (void *)doff, (void *)dcounts,
                    (void *)dpart})
      (void)hipFree(q);
  }
  return bad;
}

}  // namespace

int main() {
  try {
    std::mt19937_64 rng(7);
    int bad = check_scan(rng);
    bad += check_sort_and_bounds(rng);
    bad += check_sum(rng);
    printf(bad ? "FAILED %d\n" : "all pieces match\n", bad);
    return bad ? 1 : 0;
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
}
