// zg_peel.hip -- ZOR construction of one group on the GPU.
//
// The group's equations (one per distinct key: XOR of its H cells = value)
// are peeled in parallel rounds. A round snapshots the cells of degree 1 (the
// frontier), claims the single equation of each (one claim per equation), and
// then removes the claimed equations from all their cells; cells that drop to
// degree 1 form the next frontier. Equations claimed in the same round never
// touch each other's pivot cells (a pivot had degree 1 at the snapshot), so
// back-substitution runs one round per launch in reverse round order.
//
// Peeling to exhaustion leaves the hypergraph's 2-core, which does not depend
// on the order equations were peeled in. If a core is left, the device state
// is handed back and the host breaks it with the same rule as the CPU build
// (the caller), so both builds abandon the same keys; only the choice of
// pivot cells outside the core, and therefore the cell values, differ.
#include <mutex>

#include "zg_hash.h"
#include "zg_internal.h"

namespace zg {

__global__ void peel_init_kernel(const uint64_t *hashed, uint64_t n, uint32_t H,
                                 SegmentGeometry geo, uint32_t *slots, uint32_t *count,
                                 uint32_t *xr) {
  const uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint64_t first = hashed[e], step = splitmix64(first);
  const uint64_t base = slot_base(first, geo);
  uint64_t h = first;
  for (uint32_t i = 0; i < H; i++) {
    const uint32_t s = (uint32_t)slot_at(base, h, i, geo);
    slots[e * H + i] = s;
    atomicAdd(&count[s], 1u);
    atomicXor(&xr[s], (uint32_t)e);
    h += step;
  }
}

__global__ void peel_scan_kernel(const uint32_t *count, uint64_t S, uint32_t *frontier,
                                 uint32_t *n_frontier) {
  const uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (s < S && count[s] == 1) frontier[atomicAdd(n_frontier, 1u)] = (uint32_t)s;
}

// ctr[0]: equations peeled so far (the length of the order arrays).
__global__ void peel_claim_kernel(const uint32_t *frontier, uint32_t n_frontier,
                                  const uint32_t *count, const uint32_t *xr, uint32_t *state,
                                  uint32_t *order_e, uint32_t *order_p, uint32_t *ctr) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= n_frontier) return;
  const uint32_t s = frontier[i];
  if (count[s] != 1) return;
  const uint32_t e = xr[s];
  if (atomicCAS(&state[e], 0u, 1u) != 0u) return;
  const uint32_t k = atomicAdd(&ctr[0], 1u);
  order_e[k] = e;
  order_p[k] = s;
}

// Removes the equations claimed this round, order[begin, ctr[0]); ctr[1]
// counts the next frontier.
__global__ void peel_remove_kernel(const uint32_t *order_e, uint32_t begin, uint32_t most,
                                   const uint32_t *slots, uint32_t H, uint32_t *count,
                                   uint32_t *xr, uint32_t *next, uint32_t *ctr) {
  const uint64_t i = begin + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= (uint64_t)begin + most || i >= ctr[0]) return;
  const uint32_t e = order_e[i];
  for (uint32_t h = 0; h < H; h++) {
    const uint32_t t = slots[(uint64_t)e * H + h];
    const uint32_t old = atomicSub(&count[t], 1u);
    atomicXor(&xr[t], e);
    if (old == 2) next[atomicAdd(&ctr[1], 1u)] = t;
  }
}

// cells[pivot] = value ^ XOR of the equation's other cells, one round.
__global__ void peel_assign_kernel(const uint32_t *order_e, const uint32_t *order_p,
                                   uint32_t begin, uint32_t end, const uint32_t *slots,
                                   uint32_t H, const uint64_t *values, uint64_t *cells) {
  const uint64_t i = begin + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= end) return;
  const uint32_t e = order_e[i], p = order_p[i];
  uint64_t v = values[e];
  for (uint32_t h = 0; h < H; h++) {
    const uint32_t t = slots[(uint64_t)e * H + h];
    if (t != p) v ^= cells[t];
  }
  cells[p] = v;
}

__global__ void peel_verify_kernel(const uint32_t *slots, uint64_t n, uint32_t H,
                                   const uint64_t *values, const uint64_t *cells, uint32_t *bad) {
  const uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint64_t v = 0;
  for (uint32_t h = 0; h < H; h++) v ^= cells[slots[e * H + h]];
  if (v != values[e]) atomicAdd(bad, 1u);
}

namespace {

// Per host thread: device buffers (grown on demand) and a stream.
struct PeelWorkspace {
  DevBuf hashed, values, slots, count, xr, state, order_e, order_p, front_a, front_b, cells, ctr;
  hipStream_t stream = nullptr;
  ~PeelWorkspace() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

inline uint32_t grid(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

void device_peel(const std::vector<uint64_t> &hashed, const std::vector<uint64_t> &values,
                 uint64_t S, uint32_t H, const SegmentGeometry &geo, DevicePeel &out) {
  static thread_local PeelWorkspace ws;
  const uint64_t n = hashed.size();
  ensure(S <= UINT32_MAX && n < UINT32_MAX && (uint64_t)grid(std::max(n, S)) < (1ULL << 31),
         "group too large for the device peel", ZG_ERR_UNSUPPORTED);
  if (!ws.stream) check_hip(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking), "hipStreamCreate");
  hipStream_t s = ws.stream;
  ws.hashed.reserve(n * 8 + 8);
  ws.values.reserve(n * 8 + 8);
  ws.slots.reserve(n * H * 4 + 4);
  ws.count.reserve(S * 4);
  ws.xr.reserve(S * 4);
  ws.state.reserve(n * 4 + 4);
  ws.order_e.reserve(n * 4 + 4);
  ws.order_p.reserve(n * 4 + 4);
  ws.front_a.reserve(S * 4);
  ws.front_b.reserve(S * 4);
  ws.ctr.reserve(16);
  if (n) {
    check_hip(hipMemcpyAsync(ws.hashed.p, hashed.data(), n * 8, hipMemcpyHostToDevice, s), "copy");
    check_hip(hipMemcpyAsync(ws.values.p, values.data(), n * 8, hipMemcpyHostToDevice, s), "copy");
  }
  check_hip(hipMemsetAsync(ws.count.p, 0, S * 4, s), "memset");
  check_hip(hipMemsetAsync(ws.xr.p, 0, S * 4, s), "memset");
  check_hip(hipMemsetAsync(ws.state.p, 0, n * 4 + 4, s), "memset");
  check_hip(hipMemsetAsync(ws.ctr.p, 0, 16, s), "memset");
  uint32_t *ctr = ws.ctr.as<uint32_t>();
  if (n)
    peel_init_kernel<<<grid(n), 256, 0, s>>>(ws.hashed.as<uint64_t>(), n, H, geo,
                                            ws.slots.as<uint32_t>(), ws.count.as<uint32_t>(),
                                            ws.xr.as<uint32_t>());
  // round 0's frontier counted in ctr[1]
  peel_scan_kernel<<<grid(S), 256, 0, s>>>(ws.count.as<uint32_t>(), S, ws.front_a.as<uint32_t>(),
                                          ctr + 1);
  uint32_t h_ctr[2] = {0, 0};
  check_hip(hipMemcpyAsync(h_ctr, ctr, 8, hipMemcpyDeviceToHost, s), "copy");
  check_hip(hipStreamSynchronize(s), "device peel");
  uint32_t *front = ws.front_a.as<uint32_t>(), *next = ws.front_b.as<uint32_t>();
  uint32_t n_front = h_ctr[1], peeled = 0;
  std::vector<uint32_t> bounds{0};
  while (n_front) {
    check_hip(hipMemsetAsync(ctr + 1, 0, 4, s), "memset");
    peel_claim_kernel<<<grid(n_front), 256, 0, s>>>(front, n_front, ws.count.as<uint32_t>(),
                                                   ws.xr.as<uint32_t>(), ws.state.as<uint32_t>(),
                                                   ws.order_e.as<uint32_t>(),
                                                   ws.order_p.as<uint32_t>(), ctr);
    peel_remove_kernel<<<grid(n_front), 256, 0, s>>>(ws.order_e.as<uint32_t>(), peeled, n_front,
                                                    ws.slots.as<uint32_t>(), H,
                                                    ws.count.as<uint32_t>(), ws.xr.as<uint32_t>(),
                                                    next, ctr);
    check_hip(hipMemcpyAsync(h_ctr, ctr, 8, hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipStreamSynchronize(s), "device peel");
    peeled = h_ctr[0];
    bounds.push_back(peeled);
    n_front = h_ctr[1];
    std::swap(front, next);
  }
  out.rounds = (uint32_t)bounds.size() - 1;
  out.remaining = n - peeled;
  if (out.remaining == 0) {
    ws.cells.reserve(S * 8);
    check_hip(hipMemsetAsync(ws.cells.p, 0, S * 8, s), "memset");
    for (size_t r = bounds.size() - 1; r-- > 0;) {
      const uint32_t b = bounds[r], e = bounds[r + 1];
      if (e > b)
        peel_assign_kernel<<<grid(e - b), 256, 0, s>>>(
            ws.order_e.as<uint32_t>(), ws.order_p.as<uint32_t>(), b, e, ws.slots.as<uint32_t>(),
            H, ws.values.as<uint64_t>(), ws.cells.as<uint64_t>());
    }
    check_hip(hipMemsetAsync(ctr + 2, 0, 4, s), "memset");
    if (n)
      peel_verify_kernel<<<grid(n), 256, 0, s>>>(ws.slots.as<uint32_t>(), n, H,
                                                ws.values.as<uint64_t>(),
                                                ws.cells.as<uint64_t>(), ctr + 2);
    out.cells.resize(S);
    uint32_t bad = 0;
    check_hip(hipMemcpyAsync(out.cells.data(), ws.cells.p, S * 8, hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipMemcpyAsync(&bad, ctr + 2, 4, hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipStreamSynchronize(s), "device peel");
    ensure(bad == 0, "device back-substitution failed verification", ZG_ERR_DEVICE);
    out.order_e.clear();
    out.order_p.clear();
    return;
  }
  // a 2-core is left: the whole state goes back to the host
  out.slots.resize(n * H);
  out.count.resize(S);
  out.xr.resize(S);
  std::vector<uint32_t> st(n);
  out.order_e.resize(peeled);
  out.order_p.resize(peeled);
  check_hip(hipMemcpyAsync(out.slots.data(), ws.slots.p, n * H * 4, hipMemcpyDeviceToHost, s), "copy");
  check_hip(hipMemcpyAsync(out.count.data(), ws.count.p, S * 4, hipMemcpyDeviceToHost, s), "copy");
  check_hip(hipMemcpyAsync(out.xr.data(), ws.xr.p, S * 4, hipMemcpyDeviceToHost, s), "copy");
  if (n) check_hip(hipMemcpyAsync(st.data(), ws.state.p, n * 4, hipMemcpyDeviceToHost, s), "copy");
  if (peeled) {
    check_hip(hipMemcpyAsync(out.order_e.data(), ws.order_e.p, peeled * 4ULL, hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipMemcpyAsync(out.order_p.data(), ws.order_p.p, peeled * 4ULL, hipMemcpyDeviceToHost, s), "copy");
  }
  check_hip(hipStreamSynchronize(s), "device peel");
  out.state.resize(n);
  for (uint64_t e = 0; e < n; e++) out.state[e] = (uint8_t)st[e];
}

}  // namespace zg
