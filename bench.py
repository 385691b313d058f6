#!/usr/bin/env python3
"""Benchmark of the zorindex query hot path on one MI355X.

Workload (default): the batched query at scale -- 1024 colours of 5 Mbp from a
seeded substitution tree (0.5-5% substitutions per leaf spread over a depth-10
balanced tree, 1% indels of length 1-10 per leaf), indexed as dense canonical
31-mers with H=4 hashes and b=8 fingerprints (one 8 GiB bank of 2^23 rows x
1024 bytes). Query: 1,000,000 reads with lognormal lengths (median 2 kbp,
sigma 1.0, clipped to 100 bp-100 kbp) at 2% substitution error, random strand.
All reads sit in HBM as ASCII; one step packs them to 2 bits, extracts
k-mers and runs the lookup over every read (the autotuned order: the sorted
sweep -- per chunk of reads, radix passes by PositionHash segment, the staged
lookup from LDS segment windows, per-read sums of hit rows -- or the
persistent read-ordered lookup under a HIP graph), producing per-read
per-colour counts.

Printed: one JSON line with the throughput (queried k-mers/s), the dominant
kernel's roofline against HBM bandwidth, a bounded serving leg (Poisson
requests through the continuous-batching server: p50/p99 latency), the
one-launch-per-read path on a sample of reads, and the reference's optimised
CPU lookup restated in C, timed on the host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--colors", type=int, default=1024)
    p.add_argument("--color-len", type=int, default=5_000_000)
    p.add_argument("--reads", type=int, default=1_000_000)
    p.add_argument("--median", type=float, default=2000.0)
    p.add_argument("--sigma", type=float, default=1.0)
    p.add_argument("--min-len", type=int, default=100)
    p.add_argument("--max-len", type=int, default=100_000)
    p.add_argument("--error", type=float, default=0.02)
    p.add_argument("--bits", type=int, default=8)
    p.add_argument("--hashes", type=int, default=4)
    p.add_argument("--k", type=int, default=31)
    p.add_argument("--seed", type=int, default=2026)
    p.add_argument("--aggregate", action="store_true", help="aggregate counts only")
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--tuning", default="",
                   help="span,tiles_per_wave,prefetch,waves_per_cu[,nontemporal[,sweep]]: "
                        "fixed launch shape (skips the autotuner)")
    p.add_argument("--tune-reads", type=int, default=100000)
    p.add_argument("--plain-reads", type=int, default=2000)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-graph", action="store_true",
                   help="skip the steps with the sweep captured as one HIP graph")
    p.add_argument("--parity-reads", type=int, default=2000)
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="first copy this many MiB device-to-device (a plain copy kernel of "
                        "known size, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    p.add_argument("--mode", choices=("kmers", "minimizers"), default="kmers")
    p.add_argument("--m", type=int, default=31, help="minimizer size in minimizer mode")
    p.add_argument("--preset", choices=("batched", "minimizers", "serving", "sweep", "small"),
                   default="batched",
                   help="batched: 1024 colours x 5 Mbp, 1M lognormal reads (default); "
                        "minimizers: 4096 colours x 2 Mbp, OCM k=49 m=31, b=16, H=3, "
                        "200k reads of 10 kbp at 5%% error; serving: 8192 colours x 1 Mbp, "
                        "Poisson requests of U{1..10000} reads (median 1 kbp) through the "
                        "continuous-batching server and the plain per-request path; "
                        "small: 10 colours x 2 Mbp, 64 reads of 1 kbp")
    p.add_argument("--requests", type=int, default=300, help="serving: requests per pass")
    p.add_argument("--max-request-reads", type=int, default=10000)
    p.add_argument("--qps", default="8,16,20",
                   help="serving: Poisson arrival rates (requests/s), comma-separated; the "
                        "first is the headline")
    p.add_argument("--slots", type=int, default=2, help="serving: batches in flight")
    p.add_argument("--serve-requests", type=int, default=240,
                   help="batched workload: requests of the bounded serving leg (0: none)")
    p.add_argument("--serve-qps", type=float, default=40.0,
                   help="batched workload: Poisson arrival rate of the serving leg")
    p.add_argument("--serve-max-reads", type=int, default=64,
                   help="batched workload: reads per serving-leg request, U{1..this}")
    p.add_argument("--gpu-peel", type=int, default=0, help="build the index with the GPU peel")
    p.add_argument("--disk-load", default="",
                   help="directory to write the index to and time loading it back from")
    p.add_argument("--gather-sweep", action="store_true",
                   help="also measure random-row gather rates at 512..8192 B rows")
    p.add_argument("--sweep", default="1:2,1:3,1:8,4:2,4:3,4:8,16:2,16:3,16:8,32:2,32:3,32:8",
                   help="sweep preset: bits:hashes configurations")
    args = p.parse_args()
    given = {a.lstrip("-").split("=")[0].replace("-", "_") for a in sys.argv[1:] if a.startswith("--")}
    preset = {
        "batched": {},
        "minimizers": dict(colors=4096, color_len=2_000_000, reads=200_000, median=10_000,
                           sigma=0.0, min_len=10_000, max_len=10_000, error=0.05, bits=16,
                           hashes=3, k=49, m=31, mode="minimizers", tune_reads=5000),
        "sweep": dict(reads=200_000, steps=2, tune_reads=20000),
        "serving": dict(colors=8192, color_len=1_000_000, median=1000, sigma=1.0, min_len=100,
                        max_len=100_000, error=0.02, bits=8, hashes=4, k=31, tune_reads=20000),
        "small": dict(colors=10, color_len=2_000_000, reads=64, median=1000, sigma=0.0,
                      min_len=1000, max_len=1000, error=0.025, tune_reads=64, plain_reads=64),
    }[args.preset]
    for key, value in preset.items():
        if key not in given:
            setattr(args, key, value)
    return args


def calibration_copy(mib: int) -> None:
    """hipMemcpy device-to-device of `mib` MiB: the runtime's copy kernel reads
    and writes exactly that many bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"calibration copy: {mib} MiB device-to-device")


def make_leaves(args, synth, rng):
    """The colours: a seeded substitution tree over a random root."""
    root = synth.random_genome(rng, args.color_len)
    leaves = [None] * args.colors
    if args.colors & (args.colors - 1) == 0 and args.colors > 1:
        gen = synth.substitution_tree(rng, root, args.colors)
    else:
        gen = ((i, synth.indels(rng, synth.substitute(rng, root, rng.uniform(0.005, 0.05)), 0.01))
               for i in range(args.colors))
    for leaf, seq in gen:
        leaves[leaf] = seq
    return leaves


def build_index(args, zindex_amd, leaves, bits, hashes, device=0):
    os.environ.setdefault("ZG_BUILD_PROGRESS", "1")  # the builder reports its later phases
    t0 = time.time()
    builder = zindex_amd.Builder(k=args.k, fingerprint_bits=bits, hashes=hashes,
                                 index_mode=args.mode, minimizer_size=args.m, device=device)
    for leaf, seq in enumerate(leaves):
        builder.add_color(f"leaf{leaf:05d}", [seq.tobytes()])
        if leaf % 128 == 127:
            log(f"colours queued {leaf + 1}/{args.colors} ({time.time() - t0:.1f}s)")
    idx = builder.finish()
    log(f"index built in {time.time() - t0:.1f}s" + (" (GPU peel)" if device else ""))
    return idx


def generate(args, zindex_amd, synth):
    rng = np.random.default_rng(args.seed)
    leaves = make_leaves(args, synth, rng)
    idx = build_index(args, zindex_amd, leaves, args.bits, args.hashes, args.gpu_peel)
    lengths = synth.lognormal_lengths(rng, args.reads, args.median, args.sigma, args.min_len,
                                      args.max_len)
    t1 = time.time()
    buf, off, _ = synth.sample_reads_fast(rng, leaves, args.reads, lengths, args.error)
    log(f"{args.reads} reads, {int(off[-1])} bases sampled in {time.time() - t1:.1f}s")
    return idx, buf, off


def sweep_main(args, zindex_amd, synth) -> int:
    """Fingerprint width x hash count on the batched-query colour set: each
    index is built with the GPU peel; per configuration the query throughput
    and roofline, false negatives on stored k-mers (recovery on) and the
    false-positive rate on absent k-mers (recovery off) against 2^-b."""
    rng = np.random.default_rng(args.seed)
    leaves = make_leaves(args, synth, rng)
    lengths = synth.lognormal_lengths(rng, args.reads, args.median, args.sigma, args.min_len,
                                      args.max_len)
    buf, off, _ = synth.sample_reads_fast(rng, leaves, args.reads, lengths, args.error)
    log(f"{args.colors} colours, {args.reads} reads ({int(off[-1])} bases)")
    # stored k-mers: error-free 2 kbp pieces of 64 colours, one piece each
    picks = np.linspace(0, args.colors - 1, 64).astype(int)
    stored = [leaves[c][1000:3000].tobytes() for c in picks]
    # absent k-mers: random sequence from a stream no colour was drawn from
    arng = np.random.default_rng(args.seed + 7919)
    absent = [synth.random_genome(arng, 2000).tobytes() for _ in range(250)]
    rows = []
    for spec in args.sweep.split(","):
        bits, hashes = (int(x) for x in spec.split(":"))
        t0 = time.time()
        idx = build_index(args, zindex_amd, leaves, bits, hashes, device=1)
        build_s = time.time() - t0
        meta = idx.meta
        idx.load()
        tuning = idx.tuning
        if not args.no_autotune:
            n_t = min(args.tune_reads, args.reads)
            with idx.batch(buf[: int(off[n_t])], off[: n_t + 1], records=False) as tb:
                tuning = tb.autotune()
        with idx.batch(buf, off, records=False) as b:
            b.run()
            lk = []
            t1 = time.perf_counter()
            for _ in range(args.steps):
                b.run()
                lk.append(b.stats()["last_lookup_ms"])
            wall = (time.perf_counter() - t1) / args.steps
            st = b.stats()
        if st["tuning"]["sweep"]:
            # the sweep's lookups: the bank once per chunk, 12 B of entry and
            # the hit row per seed, over the lookup launches' HIP-event time
            algo = st["sweep_algorithmic_bytes"]
            gbs = algo / (st["sweep_lookup_ms"] / 1000) / 1e9
            kernel = "plane_lookup_kernel" if st["sweep_plane_banks"] else "stage_lookup_kernel"
        else:
            algo = st["algorithmic_bytes"] + (int(off[-1]) + args.reads) * 3 // 8
            gbs = algo / (np.mean(lk) / 1000) / 1e9
            kernel = "lookup_kernel"
        # false negatives: every stored k-mer hits its own colour
        totals, counts, _ = idx.query_reads(stored, records=True)
        fn = int(sum(int(totals[i]) - int(counts[i, c]) for i, c in enumerate(picks)))
        # false positives: absent k-mers, filter only
        t_abs, _, agg = idx.query_reads(absent, records=False, recover=False)
        n_abs = int(t_abs.sum())
        # per colour: |hits - n p| <= 8 sqrt(n p (1-p)) + 8, the reference's tolerance
        # (tests/option_coverage_cli.rs:279-285)
        p = 2.0 ** -bits
        rates = agg / max(1, n_abs)
        bound = p + (8 * np.sqrt(n_abs * p * (1 - p)) + 8) / max(1, n_abs)
        row = {"bits": bits, "hashes": hashes, "build_s": round(build_s, 1),
               "bank_bytes": meta["payload_bytes"], "abandoned_origins": meta["abandoned_origins"],
               "kmers_per_s": st["seeds"] / wall, "lookup_ms": round(float(np.mean(lk)), 3),
               "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "achieved_gbs": round(gbs, 1),
               "bytes_per_kmer": st["algorithmic_bytes"] // max(1, st["seeds"]),
               "kernel": kernel, "sweep_phases_ms": {
                   "radix_sort": round(st["sweep_sort_ms"], 2), "lookup": round(st["sweep_lookup_ms"], 2),
                   "recovery": round(st["sweep_recovery_ms"], 2), "sum": round(st["sweep_sum_ms"], 2)}
               if st["tuning"]["sweep"] else None,
               "tuning": tuning, "false_negatives": fn,
               "stored_kmers": int(totals.sum()), "absent_kmers": n_abs,
               "fpr_mean": float(rates.mean()), "fpr_max": float(rates.max()),
               "fpr_bound": bound, "fpr_within_bound": bool(np.all(rates <= bound)),
               "absent_hits_all_colours": int(agg.sum()),
               "absent_hits_expected": n_abs * p * len(agg)}
        log(f"b={bits} H={hashes}: {row}")
        rows.append(row)
        idx.close()
        del idx
    out = {
        "metric": "queried k-mers/s at 1024 colors, % of HBM peak; p50/p99 request latency",
        "value": rows[0]["kmers_per_s"], "unit": "k-mers/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": rows[0]["lookup_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u%d" % rows[0]["bits"],
        "data": "synthetic (seeded substitution tree colours and sampled reads)",
        "config": {"workload": f"fingerprint width x hash count sweep: {args.colors} colours x "
                               f"{args.color_len} bp substitution tree, k={args.k}; "
                               f"{args.reads} reads lognormal(median {args.median:.0f}, sigma "
                               f"{args.sigma}), {args.error:.0%} error; GPU-peeled indexes; "
                               f"aggregate counts", "sweep": args.sweep},
        "sweep": rows,
        "roofline": {"bound": "hbm", "achieved": rows[0]["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": rows[0]["roofline_frac"], "traffic": None,
                     "kernel": rows[0]["kernel"]},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def workload_name(args, records: bool) -> str:
    seeds = f"k={args.k}" if args.mode == "kmers" else \
        f"open-closed mod-minimizers k={args.k} m={args.m}"
    return (f"batched query: {args.colors} colours x {args.color_len} bp substitution tree, "
            f"{seeds} H={args.hashes} b={args.bits}; {args.reads} reads lognormal(median "
            f"{args.median:.0f}, sigma {args.sigma}) clipped [{args.min_len},{args.max_len}], "
            f"{args.error:.0%} error; " + ("per-read counts" if records else "aggregate counts"))


def cpu_baseline(args, idx, buf, off, gpu_counts):
    """The CPU oracle on the host: lookup phase over a bounded sample of the
    same reads (keys extracted beforehand, as the reference's lookup_s)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import ctypes as C
    import zindex_ref as zr
    zr.oracle()
    man = idx.manifest()
    payloads = [idx.bank_payload(i, copy=False) for i in range(len(man["banks"]))]
    orc = zr.OracleIndex(man, payloads, idx.recovery())
    threads = args.cpu_threads or zr.cpu_count()
    # parity (per-read per-colour counts, bit-exact) on reads drawn across the
    # whole batch, so every chunk of a chunked run is checked: evenly spaced
    # reads plus the longest ones
    sel = parity_sample(off, args.parity_reads)
    lens = np.diff(off.astype(np.int64))[sel]
    soff = np.zeros(len(sel) + 1, np.uint64)
    soff[1:] = np.cumsum(lens)
    sbuf = np.concatenate([buf[int(off[i]):int(off[i + 1])] for i in sel]) if len(sel) else buf[:0]
    exp_totals, exp_counts = orc.query_reads(sbuf, soff, threads)
    parity = None
    if gpu_counts is not None:
        parity = bool(np.array_equal(gpu_counts[sel].astype(np.uint64), exp_counts))
    # calibrate then time the lookup phase on a sample taking ~cpu_seconds
    def keys_for(n):
        parts, koff = [], np.zeros(n + 1, np.uint64)
        for i in range(n):
            read = buf[int(off[i]):int(off[i + 1])].tobytes()
            if args.mode == "minimizers":
                k = zr.scan_ocm(read, args.k, args.m, orc.manifest["seed"])[1]
            else:
                k = zr.scan_kmers(read, args.k)
            parts.append(k)
            koff[i + 1] = koff[i] + len(k)
        return (np.concatenate(parts) if parts else np.zeros(0, np.uint64)), koff
    def timed(n):
        keys, koff = keys_for(n)
        counts = np.zeros((n, orc.n_colors), np.uint64)
        t = time.perf_counter()
        zr.oracle().zo_lookup_keys_fast(C.byref(orc.z), keys.ctypes.data_as(zr.u64p),
                                        koff.ctypes.data_as(zr.u64p), n,
                                        counts.ctypes.data_as(zr.u64p), threads)
        return time.perf_counter() - t, int(koff[-1]), counts
    n = min(200, len(off) - 1)
    dt, nk, cpu_counts = timed(n)
    while dt < args.cpu_seconds / 8 and n < len(off) - 1:
        n = min(len(off) - 1, n * 4)
        dt, nk, cpu_counts = timed(n)
    if dt < args.cpu_seconds / 2 and n < len(off) - 1:
        n = min(len(off) - 1, int(n * args.cpu_seconds / max(dt, 1e-3)))
        dt, nk, cpu_counts = timed(n)
    same = None if gpu_counts is None else \
        bool(np.array_equal(gpu_counts[:n].astype(np.uint64), cpu_counts))
    return {
        "value": nk / dt, "unit": "k-mers/s" if args.mode == "kmers" else "minimizers/s",
        "cores": threads, "kind": "port",
        "sample": f"lookup phase of the reference's optimised CPU query restated in C "
                  f"(oracle/oracle.c zo_lookup_keys_fast: per-read sort with duplicate "
                  f"folding, blocks of 16 keys, rows prefetched 4 keys ahead, AVX2 screen of "
                  f"the fingerprint plane) on {threads} threads over the first {n} reads "
                  f"({nk} k-mers) of the same batch and index, {dt:.1f}s, keys pre-extracted",
        "host": host_cpu(), "counts_equal_gpu": same,
    }, parity, sel


def host_cpu() -> dict:
    """The host the CPU leg ran on: its CPU model, logical CPUs and the ones
    this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def parity_stride(off, n_reads):
    return max(1, (len(off) - 1) // max(1, n_reads - 20))


def parity_sample(off, n_reads):
    """Indices of the parity reads: every stride-th read from 0 (about
    n_reads - 20 of them) plus the 20 longest, sorted."""
    n = len(off) - 1
    if n <= 0 or n_reads <= 0:
        return np.zeros(0, np.int64)
    lens = np.diff(off.astype(np.int64))
    return np.unique(np.concatenate([np.arange(0, n, parity_stride(off, n_reads)),
                                     np.argsort(lens, kind="stable")[-20:]]))


def sample_record(sel, off, n_reads, parity):
    stride = parity_stride(off, n_reads)
    longest = sorted(int(i) for i in sel if i % stride)
    return {"reads": int(len(sel)), "bit_exact": parity,
            "indices": {"every": stride, "from": 0, "plus": longest}}


def gather_ceiling(idx, algo_bytes):
    """GB/s of random whole-row reads at the bank's row size (zg_gather_probe),
    over as many bytes as one lookup launch reads (at most 2 TiB)."""
    man = idx.manifest()
    rb = (man["banks"][0]["row_bits"] + 7) // 8
    if rb not in (512, 1024, 2048, 4096, 8192):
        return None
    n = max(1 << 20, min(algo_bytes, 1 << 41) // rb)
    out = {"gather_row_bytes": rb, "gather_ceiling_gbs": round(idx.gather_probe(rb, n), 1)}
    if rb <= 2048:  # the lookup's own row pattern: H rows of a key in H consecutive segments
        out["key_rows_ceiling_gbs"] = round(idx.gather_probe_keys(n // man["hashes"]), 1)
    return out


def percentile(values, q):
    return float(np.percentile(np.asarray(values, dtype=np.float64), q)) if len(values) else None


def serve_pass(zindex_amd, idx, seq, off, starts, sizes, arrivals, slots, max_requests):
    """One pass of requests through a server: request i (reads starts[i] ..
    starts[i]+sizes[i]) is submitted at arrivals[i] seconds after the start.
    Returns the requests' stats, the wall time and the aggregates."""
    slot_bases = max(int(off[a + n] - off[a]) for a, n in zip(starts, sizes))
    srv = zindex_amd.Server(idx, slots=slots, max_requests=max_requests,
                            slot_reads=max(32768, int(max(sizes))),
                            slot_bases=max(64 << 20, slot_bases))
    tickets = []
    t0 = time.perf_counter()
    for a, n, at in zip(starts, sizes, arrivals):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(seq, off[a: a + n + 1], records=False, aggregate=True))
    stats, aggs = [], []
    for t in tickets:
        _, _, agg, st = srv.wait(t)
        stats.append(st)
        aggs.append(agg)
    wall = time.perf_counter() - t0
    sst = srv.stats()
    srv.close()
    return stats, wall, aggs, sst


def serving_leg(args, zindex_amd, idx, buf, off, counts):
    """A bounded serving pass over the batch's index: requests of U{1..
    serve_max_reads} consecutive reads from the start of the batch, Poisson
    arrivals at serve_qps, through the continuous-batching server (per-request
    aggregates). Latency percentiles, and each request's aggregate checked
    against the batch's per-read counts of the same reads."""
    rng = np.random.default_rng(args.seed + 3)
    sizes = rng.integers(1, args.serve_max_reads + 1, args.serve_requests)
    n = int(sizes.sum())
    if n > len(off) - 1:
        return None
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    pinned = zindex_amd.HostBuffer((int(off[n]),), np.uint8)
    pinned.array[:] = buf[: int(off[n])]
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_qps, args.serve_requests))
    serve_pass(zindex_amd, idx, pinned.array, off, starts[:4], sizes[:4], np.zeros(4), args.slots,
               256)  # warm-up
    st, wall, aggs, sst = serve_pass(zindex_amd, idx, pinned.array, off, starts, sizes, arrivals,
                                     args.slots, 256)
    pinned.free()
    lat = [x["latency_ms"] for x in st]
    seeds = sum(x["seeds"] for x in st)
    same = None
    if counts is not None:
        same = all(np.array_equal(agg, counts[a: a + k].astype(np.uint64).sum(axis=0))
                   for a, k, agg in zip(starts, sizes, aggs))
    out = {"requests": int(args.serve_requests), "qps": args.serve_qps,
           "reads_per_request": f"U{{1..{args.serve_max_reads}}}", "slots": args.slots,
           "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
           "mean_ms": float(np.mean(lat)), "max_ms": float(np.max(lat)),
           "kmers_per_s": seeds / wall, "batches": sst["batches"],
           "max_requests_per_batch": sst["max_batch_requests"], "equals_batch_counts": same}
    log(f"serving leg: {args.serve_requests} requests at {args.serve_qps}/s: p50 "
        f"{out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, {sst['batches']} batches, "
        f"aggregates equal the batch's counts: {same}")
    return out


def serve_main(args, zindex_amd, synth) -> int:
    """The serving workload: Poisson arrivals of requests at a fixed rate through
    the continuous-batching server, and the same arrivals through the plain
    per-request path (one request per batch, one batch at a time)."""
    rng = np.random.default_rng(args.seed + 1)
    sizes = rng.integers(1, args.max_request_reads + 1, args.requests)
    args.reads = int(sizes.sum())
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    idx, buf, off = generate(args, zindex_amd, synth)
    t0 = time.time()
    idx.load()
    meta = idx.meta
    log(f"index resident in HBM ({meta['payload_bytes'] / 2**30:.2f} GiB) in {time.time() - t0:.1f}s")
    if args.tuning:
        vals = [int(x) for x in args.tuning.split(",")]
        idx.tuning = {"span": vals[0], "tiles_per_wave": vals[1], "prefetch": vals[2],
                      "waves_per_cu": vals[3], "nontemporal": vals[4] if len(vals) > 4 else 0}
    elif not args.no_autotune:
        n_t = min(args.tune_reads, args.reads)
        with idx.batch(buf[: int(off[n_t])], off[: n_t + 1], records=True) as tb:
            log(f"autotuned on {n_t} reads: {tb.autotune()}")
    # the requests' reads in page-locked memory (DMA straight from the caller)
    pinned = zindex_amd.HostBuffer((len(buf),), np.uint8)
    pinned.array[:] = buf
    seq = pinned.array
    man = idx.manifest()
    per_lookup = sum(meta["hashes"] * ((len(b["groups"]) * meta["fingerprint_bits"] + 7) // 8)
                     // ((b.get("routing") or {}).get("partitions", 1)) for b in man["banks"])
    # warm-up: a few requests through each server shape
    for slots, mr in ((args.slots, 256), (1, 1)):
        serve_pass(zindex_amd, idx, seq, off, starts[:4], sizes[:4], np.zeros(4), slots, mr)
    results = {}
    zero = np.zeros(args.requests)
    for name, slots, mr in (("batched", args.slots, 256), ("plain", 1, 1)):
        st, wall, aggs, sst = serve_pass(zindex_amd, idx, seq, off, starts, sizes, zero, slots, mr)
        seeds = sum(x["seeds"] for x in st)
        results[name] = {"saturation_kmers_per_s": seeds / wall, "wall_s": wall,
                         "seeds": seeds, "batches": sst["batches"],
                         "mean_requests_per_batch": args.requests / max(1, sst["batches"]),
                         "lookup_ms_sum": sst["lookup_ms"],
                         "sweep_batches": sst["sweep_batches"],
                         "device_bytes": sst["device_bytes"]}
        results[name + "_aggs"] = aggs
        log(f"{name} closed loop: {seeds / wall:.3e} k-mers/s, {sst['batches']} batches "
            f"({sst['sweep_batches']} as the sorted sweep), {wall:.2f}s")
    batched_aggs, plain_aggs = results.pop("batched_aggs"), results.pop("plain_aggs")
    same = all(np.array_equal(a, b) for a, b in zip(batched_aggs, plain_aggs))
    # requests alone through the batch query, on a sample, against the mixed pass
    alone_same = True
    for i in [int(i) for i in np.linspace(0, args.requests - 1, 6)]:
        a, n = int(starts[i]), int(sizes[i])
        _, _, agg = idx.query_reads(buf[int(off[a]): int(off[a + n])], off[a: a + n + 1] - off[a])
        alone_same &= bool(np.array_equal(agg, batched_aggs[i]))
    log(f"mixed == plain: {same}; alone == mixed on 6 requests: {alone_same}")
    sweep = []
    for q in [float(x) for x in args.qps.split(",")]:
        arrivals = np.cumsum(np.random.default_rng(args.seed + 2).exponential(1.0 / q, args.requests))
        row = {"qps": q}
        for name, slots, mr in (("batched", args.slots, 256), ("plain", 1, 1)):
            st, wall, aggs, sst = serve_pass(zindex_amd, idx, seq, off, starts, sizes, arrivals,
                                             slots, mr)
            lat = [x["latency_ms"] for x in st]
            seeds = sum(x["seeds"] for x in st)
            row[name] = {"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                         "mean_ms": float(np.mean(lat)), "kmers_per_s": seeds / wall,
                         "mean_requests_per_batch": args.requests / max(1, sst["batches"]),
                         "max_requests_per_batch": sst["max_batch_requests"],
                         "sweep_batches": sst["sweep_batches"]}
            log(f"qps {q} {name}: p50 {row[name]['p50_ms']:.1f} ms, p99 {row[name]['p99_ms']:.1f} ms, "
                f"{seeds / wall:.3e} k-mers/s")
        sweep.append(row)
    b = results["batched"]
    # the read-ordered lookup kernel alone at this geometry: one persistent
    # launch over the first requests' reads (aggregate counts), timed with HIP
    # events
    n_k = int(min(args.reads, 200_000))
    served_tuning = dict(idx.tuning)
    idx.tuning = dict(served_tuning, sweep=0)
    with idx.batch(buf[: int(off[n_k])], off[: n_k + 1], records=False) as kb:
        kb.run()
        lk = []
        for _ in range(3):
            kb.run()
            lk.append(kb.stats()["last_lookup_ms"])
        kst = kb.stats()
    idx.tuning = served_tuning
    positions = int(off[n_k]) + n_k
    algo = kst["algorithmic_bytes"] + positions * 3 // 8
    lk_ms = float(np.mean(lk))
    kernel_gbs = algo / (lk_ms / 1000) / 1e9
    log(f"lookup kernel on {n_k} reads: {lk_ms:.1f} ms, {kernel_gbs:.0f} GB/s")
    ceiling = gather_ceiling(idx, algo)
    cpu = None
    if not args.no_cpu:
        cpu, _, _ = cpu_baseline(args, idx, buf, off, None)
        log(f"cpu baseline: {cpu['value']:.3e} k-mers/s on {cpu['cores']} threads")
    pinned.free()
    head = sweep[0]
    out = {
        "metric": "queried k-mers/s at 1024 colors, % of HBM peak; p50/p99 request latency",
        "value": b["saturation_kmers_per_s"],
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": 4,
        "ms_per_step": b["wall_s"] * 1000 / args.requests,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if args.bits == 8 else f"u{args.bits}",
        "data": "synthetic (seeded substitution tree colours and sampled reads)",
        "config": {
            "workload": (f"serving: {args.colors} colours x {args.color_len} bp substitution tree, "
                         f"k={args.k} H={args.hashes} b={args.bits}; {args.requests} requests of "
                         f"U{{1..{args.max_request_reads}}} reads, lognormal(median "
                         f"{args.median:.0f}, sigma {args.sigma}) clipped [{args.min_len},"
                         f"{args.max_len}], {args.error:.0%} error; Poisson arrivals at "
                         f"{head['qps']} requests/s; per-request aggregate counts"),
            "colors": args.colors, "reads": args.reads, "bases": int(off[-1]),
            "bank_bytes": meta["payload_bytes"], "slots": args.slots, "tuning": idx.tuning,
            "step": "one request: H2D from page-locked memory, pack, k-mers, the lookups (the "
                    "persistent read-ordered lookup, or the sorted sweep for batches the server's "
                    "model finds faster that way), per-request sums, D2H",
        },
        "latency": {"qps": head["qps"], "batched_p50_ms": head["batched"]["p50_ms"],
                    "batched_p99_ms": head["batched"]["p99_ms"],
                    "plain_p50_ms": head["plain"]["p50_ms"], "plain_p99_ms": head["plain"]["p99_ms"]},
        "qps_sweep": sweep,
        "saturation": {k: results[k] for k in ("batched", "plain")},
        "roofline": {"bound": "hbm", "achieved": round(kernel_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(kernel_gbs / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "lookup_kernel", "launch_ms": round(lk_ms, 3),
                     "algorithmic_bytes_per_launch": algo,
                     "launch": f"one persistent launch over the first {n_k} reads, aggregate "
                               f"counts, recovery map on", **(ceiling or {}),
                     "frac_of_gather_ceiling": round(kernel_gbs / ceiling["gather_ceiling_gbs"], 4)
                     if ceiling else None},
        "closed_loop_row_gbs": round(b["seeds"] * per_lookup / b["wall_s"] / 1e9, 1),
        "cpu_baseline": cpu,
        "mixed_equals_plain": same,
        "alone_equals_mixed": alone_same,
    }
    print(json.dumps(out))
    return 0


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the query runs on one GPU; --gpus must be 1 (sharding "
                                   "banks or reads over several GPUs is designed, not built)"}))
        return 2
    import zindex_amd
    from zindex_amd import synth
    if args.preset == "serving":
        return serve_main(args, zindex_amd, synth)
    if args.preset == "sweep":
        return sweep_main(args, zindex_amd, synth)

    records = not args.aggregate
    if args.calibrate_copy_mib:
        calibration_copy(args.calibrate_copy_mib)
    idx, buf, off = generate(args, zindex_amd, synth)
    meta = idx.meta
    t0 = time.time()
    idx.load()
    load_s = time.time() - t0
    log(f"index resident in HBM ({meta['payload_bytes'] / 2**30:.2f} GiB) in {load_s:.1f}s")
    load = {"bank_bytes": meta["payload_bytes"], "from_host_s": round(load_s, 3)}
    if args.disk_load:
        # the reference's load phase: bank files -> memory (here, decoded
        # frame by frame into page-locked chunks copied to HBM)
        import tempfile
        with tempfile.TemporaryDirectory(dir=args.disk_load) as d:
            t0 = time.time()
            idx.write(os.path.join(d, "idx"))
            load["write_s"] = round(time.time() - t0, 3)
            load["file_bytes"] = sum(f.stat().st_size for f in Path(d, "idx").iterdir())
            t0 = time.time()
            with zindex_amd.Index.open(os.path.join(d, "idx")) as z:
                z.load()
                load["from_disk_s"] = round(time.time() - t0, 3)
        load["from_disk_gbs"] = round(meta["payload_bytes"] / load["from_disk_s"] / 1e9, 2)
        log(f"load: {load}")

    tuning = idx.tuning
    if args.tuning:
        vals = [int(x) for x in args.tuning.split(",")]
        span, tiles, prefetch, waves = vals[:4]
        nontemporal = vals[4] if len(vals) > 4 else 0
        idx.tuning = {"span": span, "tiles_per_wave": tiles, "prefetch": prefetch,
                      "waves_per_cu": waves, "nontemporal": nontemporal,
                      "sweep": vals[5] if len(vals) > 5 else 0}
        tuning = idx.tuning
    elif not args.no_autotune:
        n_t = min(args.tune_reads, args.reads)
        with idx.batch(buf[: int(off[n_t])], off[: n_t + 1], records=records) as tb:
            tuning = tb.autotune()
        log(f"autotuned on {n_t} reads: {tuning}")

    t0 = time.time()
    batch = idx.batch(buf, off, records=records)
    log(f"batch resident + seeded in {time.time() - t0:.1f}s")
    for _ in range(args.warmup):
        batch.reseed()
        batch.run()
    t0 = time.perf_counter()
    lookup_ms = []
    for _ in range(args.steps):
        batch.reseed()
        batch.run()
        lookup_ms.append(batch.stats()["last_lookup_ms"])
    elapsed = time.perf_counter() - t0
    st = batch.stats()
    seeds = st["seeds"]
    ms_per_step = elapsed * 1000 / args.steps
    value = seeds * args.steps / elapsed
    totals, counts, agg = batch.results()
    log(f"{seeds} k-mers per step, {ms_per_step:.1f} ms/step, lookup {np.mean(lookup_ms):.1f} ms")
    if st["tuning"]["sweep"]:
        log(f"sweep phases (last step): sort {st['sweep_sort_ms']:.1f} ms, lookup "
            f"{st['sweep_lookup_ms']:.1f} ms, recovery {st['sweep_recovery_ms']:.1f} ms, "
            f"sum {st['sweep_sum_ms']:.1f} ms, "
            f"{st['sweep_launches']} chunks, {st['sweep_folded']} folded; seeding "
            f"{st['last_seed_ms']:.1f} ms")

    # the same steps with the sweep captured as one HIP graph per batch (its
    # phases are then not timed, so the line's value and roofline come from
    # the eager steps above)
    graph = None
    if st["tuning"]["sweep"] and not args.no_graph:
        tuned = dict(idx.tuning)
        idx.tuning = dict(tuned, graph=1)
        batch.reseed()
        batch.run()  # eager, then captured
        g0 = time.perf_counter()
        for _ in range(args.steps):
            batch.reseed()
            batch.run()
        g_elapsed = time.perf_counter() - g0
        gst = batch.stats()
        _, g_counts, g_agg = batch.results()
        graph = {"captured": bool(gst["tuning"]["graph"]),
                 "ms_per_step": g_elapsed * 1000 / args.steps,
                 "kmers_per_s": seeds * args.steps / g_elapsed,
                 "equals_eager": bool(np.array_equal(g_agg, agg) and
                                      (g_counts is None or np.array_equal(g_counts, counts)))}
        idx.tuning = tuned
        log(f"graph steps: {graph}")

    # roofline of the dominant kernel: algorithmic bytes per launch / its
    # average launch time (HIP events on the stream the kernel runs on)
    positions = int(off[-1]) + args.reads
    lk_ms = float(np.mean(lookup_ms))
    sweep = bool(st["tuning"]["sweep"])
    if sweep:
        # sorted sweep: the staged lookup, one launch per chunk of reads and
        # bank, reads every bank row once, each seed's 8-byte first hash and
        # 4-byte position once, and writes each seed's hit row once
        launches = max(1, st["sweep_stage_launches"])
        algo = st["sweep_algorithmic_bytes"] // launches
        kernel_ms = st["sweep_lookup_ms"] / launches
        kernel = "plane_lookup_kernel" if st["sweep_plane_banks"] else "stage_lookup_kernel"
    elif args.mode == "kmers":  # packed bases and invalid bits read by the lookup
        algo = st["algorithmic_bytes"] + positions * 3 // 8
        kernel_ms = lk_ms
        kernel = "lookup_kernel"
    else:  # explicit minimizer keys
        algo = st["algorithmic_bytes"] + seeds * 8
        kernel_ms = lk_ms
        kernel = "lookup_kernel"
    achieved = algo / (kernel_ms / 1000) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": kernel, "launch_ms": round(kernel_ms, 3),
                "algorithmic_bytes_per_launch": algo}
    if sweep:
        sum_bytes = seeds * st["sweep_hit_row_bytes"]  # hit rows read back once
        roofline.update({
            "launches_per_step": launches, "chunks_per_step": st["sweep_launches"],
            "seeds_per_step": seeds, "hit_row_bytes": st["sweep_hit_row_bytes"],
            "folded_per_step": st["sweep_folded"],
            "phases_ms": {"seeding": round(st["last_seed_ms"], 2),
                          "radix_sort": round(st["sweep_sort_ms"], 2),
                          "lookup": round(st["sweep_lookup_ms"], 2),
                          "recovery": round(st["sweep_recovery_ms"], 2),
                          "sum": round(st["sweep_sum_ms"], 2)},
            "sum_kernel_gbs": round(sum_bytes / (st["sweep_sum_ms"] / 1000) / 1e9, 1)
            if st["sweep_sum_ms"] > 0 else None})
        # the lookups' hit-row stores against the same rows stored alone (one
        # whole row per seed of the last chunk, in the sweep's order, each row
        # written by neighbouring lanes of one store instruction): the store
        # ceiling of the lookups' position pattern
        try:
            ceiling = batch.store_probe()
        except zindex_amd._lib.ZindexError as e:
            log(f"store probe: {e}")
            ceiling = None
        if ceiling:
            write_gbs = sum_bytes / launches / (kernel_ms / 1000) / 1e9
            roofline.update({"hit_row_write_gbs": round(write_gbs, 1),
                             "store_replay_gbs": round(ceiling, 1),
                             "frac_of_store_replay": round(write_gbs / ceiling, 4)})
            log(f"hit-row stores {write_gbs:.0f} GB/s; replayed alone {ceiling:.0f} GB/s "
                f"({write_gbs / ceiling:.1%})")
    else:
        # the ceiling of the row gathers alone: random whole rows of the
        # bank's row size, the same bytes per launch, no other work
        ceiling = gather_ceiling(idx, algo)
        if ceiling:
            roofline.update(ceiling)
            roofline["frac_of_gather_ceiling"] = round(achieved / ceiling["gather_ceiling_gbs"], 4)
            if "key_rows_ceiling_gbs" in ceiling:
                roofline["frac_of_key_rows_ceiling"] = round(
                    achieved / ceiling["key_rows_ceiling_gbs"], 4)
            log(f"gather ceiling at {ceiling['gather_row_bytes']} B rows: "
                f"{ceiling['gather_ceiling_gbs']:.0f} GB/s (key-row pattern "
                f"{ceiling.get('key_rows_ceiling_gbs')}); lookup at "
                f"{roofline['frac_of_gather_ceiling']:.1%} of it")
    if args.gather_sweep:
        roofline["gather_sweep_gbs"] = {rb: round(idx.gather_probe(rb, (1 << 40) // rb), 1)
                                        for rb in (512, 1024, 2048, 4096, 8192)}
        log(f"gather sweep: {roofline['gather_sweep_gbs']}")
    log(f"roofline: {kernel} {achieved:.1f} GB/s algorithmic over {kernel_ms:.2f} ms per launch "
        f"= {achieved / HBM_PEAK_GBS:.1%} of HBM peak")
    workload = workload_name(args, records)
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE passes of this
    # workload (scaled by a copy kernel of known size), and the kernel's
    # instruction mix from SQ counters, when recorded for it
    tuned = ",".join(str(st["tuning"][k]) for k in ("span", "tiles_per_wave", "prefetch",
                                                      "waves_per_cu", "nontemporal", "sweep"))
    for tf in sorted((ROOT / "profiles").glob("*.json")):
        rec = json.loads(tf.read_text())
        if not (rec.get("workload") == workload and rec.get("kernel_name") == kernel
                and rec.get("algorithmic_bytes") == algo
                and tuned.startswith(rec.get("tuning", ""))):
            continue
        if "traffic_bytes" in rec:
            roofline["traffic"] = rec["traffic_bytes"]
            roofline["traffic_source"] = str(tf.relative_to(ROOT))
            # the measured HBM bytes at this run's launch time
            roofline["traffic_gbs"] = round(rec["traffic_bytes"] / (kernel_ms / 1000) / 1e9, 1)
            roofline["traffic_frac"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 4)
        if "issue" in rec:
            roofline["issue"] = dict(rec["issue"], source=str(tf.relative_to(ROOT)))

    plain = None
    if args.plain_reads:
        n_p = min(args.plain_reads, args.reads)
        with idx.batch(buf[: int(off[n_p])], off[: n_p + 1], records=records) as pb:
            pb.run_plain()
            pst = pb.stats()
            pt, pc, _ = pb.results()
            pb.run()
            bst = pb.stats()
            bt, bc, _ = pb.results()
            same = bool(np.array_equal(pc, bc)) if records else None
        plain = {"reads": n_p, "kmers": pst["seeds"],
                 "plain_kmers_per_s": pst["seeds"] / (pst["last_total_ms"] / 1000),
                 "batched_kmers_per_s_same_reads": bst["seeds"] / (bst["last_total_ms"] / 1000),
                 "plain_equals_batched": same}
        log(f"plain path: {plain}")

    cpu, parity, sel = (None, None, np.zeros(0, np.int64))
    if not args.no_cpu:
        cpu, parity, sel = cpu_baseline(args, idx, buf, off, counts)
        log(f"cpu baseline: {cpu['value']:.3e} k-mers/s on {cpu['cores']} threads; "
            f"parity on {len(sel)} reads across the batch: {parity}")

    latency = None
    if args.serve_requests > 0:
        latency = serving_leg(args, zindex_amd, idx, buf, off, counts if records else None)

    out = {
        "metric": "queried k-mers/s at 1024 colors, % of HBM peak; p50/p99 request latency",
        "value": value,
        "unit": "k-mers/s" if args.mode == "kmers" else "minimizers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if args.bits == 8 else f"u{args.bits}",
        "data": "synthetic (seeded substitution tree colours and sampled reads)",
        "config": {
            "workload": workload,
            "colors": args.colors, "k": args.k, "hashes": args.hashes, "bits": args.bits,
            "reads": args.reads, "bases": int(off[-1]), "kmers_per_step": seeds,
            "bank_bytes": meta["payload_bytes"], "tuning": st["tuning"],
            "step": ("pack ASCII->2-bit + k-mer counts" if args.mode == "kmers" else
                     "pack ASCII->2-bit + open-closed minimizer selection and compaction")
                    + (f" + per chunk of whole reads ({st['sweep_launches']} chunks): stable LSD "
                       "radix passes of the seeds by PositionHash segment (6-bit digits), the "
                       + ("bit-plane lookup" if st["sweep_plane_banks"] else "staged lookup")
                       + " from LDS segment windows into hit rows, per-read sums "
                       "of the hit rows; eager launches with no host synchronization inside "
                       "the step" if sweep else " + persistent lookup (HIP graph)"),
        },
        "roofline": roofline,
        "graph": graph,
        "latency": latency,
        "cpu_baseline": cpu,
        "plain_path": plain,
        "speedup_vs_plain": (plain["batched_kmers_per_s_same_reads"] / plain["plain_kmers_per_s"])
        if plain else None,
        "parity_sample": sample_record(sel, off, args.parity_reads, parity)
        if parity is not None else None,
        "load": load,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
