// zg_scan.hip -- device-wide exclusive prefix sums for gfx950.
//
// Reduce-then-scan over tiles of 4096 elements (256 threads x 16): one
// kernel writes each tile's sum, the tile sums are scanned the same way
// (recursively, down to a single tile), and a second kernel scans every tile
// from its scanned block offset. No block waits on another, so the scan has
// no inter-block spin and is safe to capture in a HIP graph. Traffic: the
// input is read twice and the output written once.
#include <hip/hip_runtime.h>

#include "zg_device.h"
#include "zg_format.h"
#include "zg_kernels.h"

namespace zg {

namespace {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 16;
constexpr uint64_t kScanTile = kScanThreads * kScanItems;

template <class T>
__global__ __launch_bounds__(256) void scan_reduce_kernel(const T *in, uint64_t n, uint64_t *sums) {
  __shared__ uint64_t sh[4];
  const uint64_t base = blockIdx.x * kScanTile;
  uint64_t s = 0;
#pragma unroll
  for (uint32_t r = 0; r < kScanItems; r++) {
    const uint64_t i = base + r * kScanThreads + threadIdx.x;  // coalesced
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t total;
  (void)block_exclusive_scan<4>(s, sh, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Thread t owns elements base + 16t .. base + 16t + 15 of its tile; every
// thread reads its elements before any thread writes, so in == out works.
template <class T>
__global__ __launch_bounds__(256) void scan_down_kernel(const T *in, uint64_t n,
                                                        const uint64_t *block_off, uint64_t *out) {
  __shared__ uint64_t sh[4];
  const uint64_t base = blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
#pragma unroll
  for (uint32_t j = 0; j < kScanItems; j++) v[j] = base + j < n ? (uint64_t)in[base + j] : 0;
  uint64_t run = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanItems; j++) {
    const uint64_t x = v[j];
    v[j] = run;
    run += x;
  }
  uint64_t total;
  const uint64_t before = block_exclusive_scan<4>(run, sh, total);
  const uint64_t off = (block_off ? block_off[blockIdx.x] : 0) + before;
#pragma unroll
  for (uint32_t j = 0; j < kScanItems; j++)
    if (base + j < n) out[base + j] = v[j] + off;
}

uint64_t aligned_bytes(uint64_t words) { return (words * 8 + 255) & ~255ULL; }

uint64_t temp_bytes(uint64_t n) {
  uint64_t bytes = 256;
  while (n > kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    bytes += aligned_bytes(n);
  }
  return bytes;
}

template <class T>
void scan_levels(const T *in, uint64_t *out, uint64_t n, uint8_t *temp, hipStream_t s) {
  if (!n) return;
  const uint64_t blocks = (n + kScanTile - 1) / kScanTile;
  if (blocks == 1) {
    scan_down_kernel<T><<<1, kScanThreads, 0, s>>>(in, n, nullptr, out);
    return;
  }
  uint64_t *sums = (uint64_t *)temp;
  scan_reduce_kernel<T><<<(uint32_t)blocks, kScanThreads, 0, s>>>(in, n, sums);
  scan_levels<uint64_t>(sums, sums, blocks, temp + aligned_bytes(blocks), s);
  scan_down_kernel<T><<<(uint32_t)blocks, kScanThreads, 0, s>>>(in, n, sums, out);
}

}  // namespace

size_t scan_temp_bytes_u64(uint64_t n) { return temp_bytes(n); }
size_t scan_temp_bytes_u32_to_u64(uint64_t n) { return temp_bytes(n); }

void launch_exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, void *temp,
                               size_t temp_bytes_given, hipStream_t s) {
  ensure(temp_bytes_given >= temp_bytes(n), "scan: temporary buffer too small", -4);
  scan_levels<uint64_t>(in, out, n, (uint8_t *)temp, s);
}

void launch_exclusive_scan_u32_to_u64(const uint32_t *in, uint64_t *out, uint64_t n, void *temp,
                                      size_t temp_bytes_given, hipStream_t s) {
  ensure(temp_bytes_given >= temp_bytes(n), "scan: temporary buffer too small", -4);
  scan_levels<uint32_t>(in, out, n, (uint8_t *)temp, s);
}

}  // namespace zg
