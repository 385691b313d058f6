"""GPU parity at the shapes the benchmark configurations run: recovery keys
with more postings than one wave holds, the sorted sweep at the batched
query's launch shape over 1024+ colours (one chunk and many), and the small
plain-method configuration at its stated size. Every case is bit-exact
against the CPU oracle on per-read totals, per-read per-colour counts and the
aggregate."""
import numpy as np
import pytest

import zindex_ref as zr

zindex_amd = pytest.importorskip("zindex_amd")
from zindex_amd import synth  # noqa: E402
from test_gpu_parity import (assert_sweep_runs, build, check_parity, make_colors,  # noqa: E402
                             make_reads, oracle_for, say)

pytestmark = pytest.mark.gpu

# the batched query's autotuned shape: sorted sweep, 64-seed spans, one tile,
# one seed in flight, 20 waves per CU
SWEEP_BENCH = {"span": 64, "tiles_per_wave": 1, "prefetch": 1, "waves_per_cu": 20,
               "nontemporal": 0, "sweep": 1}
READ_ORDER = {"span": 16384, "tiles_per_wave": 1, "prefetch": 2, "waves_per_cu": 0,
              "nontemporal": 1, "sweep": 0}


def postings_of_query_keys(idx, reads, k=31):
    """Largest posting count among the recovery entries the reads' k-mers hit."""
    hashes, offs, _ = idx.recovery()
    count = dict(zip(hashes.tolist(), np.diff(offs).tolist()))
    seed = idx.manifest()["seed"]
    o = zr.oracle()
    best = 0
    for r in reads:
        for key in zr.scan_kmers(r, k).tolist():
            best = max(best, count.get(o.zo_abandoned_hash(key, seed), 0))
    return best


# the sweep's lookup kernels: over bit-plane copies of the banks (banks
# without merged groups), or the staged lookup over the byte planes
KERNELS = {"planes": "1", "staged": "0"}


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("group", [1, 4])
def test_recovery_keys_with_many_postings(group, kernel, monkeypatch):
    # 160 near-identical colours at load 2 (slot_scale 0.5): k-mers shared by
    # every colour stay in the 2-core of every column, so their recovery
    # entries list well over 64 colours and the lookup walks their postings
    # in several passes of 64 lanes (storage.rs:2771-2787, 2916-2924)
    colors = make_colors(300 + group, 160, 5000, rates=(0.0, 0.003))
    reads = make_reads(301 + group, colors, 120, lo=20, hi=2000)
    idx = build(colors, fingerprint_bits=8, hashes=4, slot_scale=0.5, merge_group_size=group)
    most = postings_of_query_keys(idx, reads)
    say(f"group {group}: largest posting list hit by the reads: {most}")
    assert most > 128
    monkeypatch.setenv("ZG_SWEEP_PLANES", KERNELS[kernel])
    idx.load(recover=True)
    orc = oracle_for(idx, True)
    for name, tuning in (("read order", READ_ORDER), ("sweep", SWEEP_BENCH)):
        idx.tuning = tuning
        check_parity(idx, orc, reads, True, label=f"postings>128 g{group} {name}")
    st = assert_sweep_runs(idx, reads)
    assert st["sweep_plane_banks"] == (1 if kernel == "planes" and group == 1 else 0), st
    monkeypatch.setenv("ZG_SWEEP_CHUNK_POSITIONS", "20000")
    check_parity(idx, orc, reads, True, label=f"postings>128 g{group} {kernel} sweep chunks")


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("n_colors", [1024, 1100])
def test_sorted_sweep_at_bench_shape(n_colors, kernel, monkeypatch):
    # 1024 singleton colours (one 1 KiB tile per row) and 1100 (two tiles)
    rng = np.random.default_rng(400 + n_colors)
    root = synth.random_genome(rng, 4000)
    colors = [synth.substitute(rng, root, float(rng.uniform(0.005, 0.05))).tobytes()
              for _ in range(n_colors)]
    reads = make_reads(401, colors, 300, lo=30, hi=3000)
    idx = build(colors, fingerprint_bits=8, hashes=4)
    monkeypatch.setenv("ZG_SWEEP_PLANES", KERNELS[kernel])
    idx.load()
    orc = oracle_for(idx)
    idx.tuning = SWEEP_BENCH
    st = assert_sweep_runs(idx, reads)
    assert st["sweep_plane_banks"] == (1 if kernel == "planes" else 0), st
    check_parity(idx, orc, reads, label=f"sweep bench shape {n_colors} colours {kernel}")
    # many chunks: each a few reads, chunk starts far from position 0
    monkeypatch.setenv("ZG_SWEEP_CHUNK_POSITIONS", "7000")
    check_parity(idx, orc, reads, label=f"sweep bench shape {n_colors} colours {kernel}, chunks")


def test_plain_method_small_at_stated_size():
    # 10 colours x 2 Mbp: colour 0 uniform i.i.d. ACGT, colours 1-9 from it at
    # U[1%, 20%] substitutions; k=31 H=4 b=8. Query: 64 fragments of 1 kbp
    # from random colours and offsets, half reverse-complemented, U[0, 5%]
    # substitutions each
    rng = np.random.default_rng(2024)
    root = synth.random_genome(rng, 2_000_000)
    colors = [root] + [synth.substitute(rng, root, float(rng.uniform(0.01, 0.20)))
                       for _ in range(9)]
    reads = []
    for _ in range(64):
        c = colors[int(rng.integers(0, 10))]
        start = int(rng.integers(0, len(c) - 1000 + 1))
        piece = synth.substitute(rng, c[start:start + 1000], float(rng.uniform(0.0, 0.05)))
        if rng.random() < 0.5:
            piece = synth.reverse_complement(piece)
        reads.append(piece.tobytes())
    idx = build([c.tobytes() for c in colors], fingerprint_bits=8, hashes=4)
    idx.load()
    orc = oracle_for(idx)
    for name, tuning in (("default", idx.tuning), ("read order", READ_ORDER),
                         ("sweep", SWEEP_BENCH)):
        idx.tuning = tuning
        counts = check_parity(idx, orc, reads, label=f"10 x 2 Mbp, 64 x 1 kbp, {name}")
    assert counts.sum() > 0


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("case", ["h4 recovery", "h3 partitioned", "h2"])
def test_lane_lookup_wide_rows(case, kernel, monkeypatch):
    # rows of >= 512 colours at b = 8 (whole 512-byte slices, one entry per
    # lane in the staged lookup), mixed with the wave-per-entry pass for
    # entries with recovery postings, routed to partitions, and a last slice
    # holding fewer than 512 fields
    rng = np.random.default_rng(500 + len(case))
    if case == "h4 recovery":
        colors = make_colors(501, 600, 3000, rates=(0.0, 0.003))
        cfg = dict(hashes=4, slot_scale=0.5)
    elif case == "h3 partitioned":
        colors = make_colors(502, 560, 3000)
        cfg = dict(hashes=3, slot_scale=0.8, max_subindex_size=1 << 20)
    else:
        colors = make_colors(503, 530, 2500)
        cfg = dict(hashes=2)
    reads = make_reads(int(rng.integers(1 << 30)), colors, 200, lo=30, hi=2500)
    idx = build(colors, fingerprint_bits=8, **cfg)
    if case == "h3 partitioned":
        assert idx.meta["n_banks"] > 1
    if case == "h4 recovery":
        most = postings_of_query_keys(idx, reads)
        say(f"{case}: largest posting list hit by the reads: {most}")
        assert most > 64
    monkeypatch.setenv("ZG_SWEEP_PLANES", KERNELS[kernel])
    idx.load(recover=True)
    orc = oracle_for(idx, True)
    idx.tuning = SWEEP_BENCH
    st = assert_sweep_runs(idx, reads)
    assert (st["sweep_plane_banks"] > 0) == (kernel == "planes"), st
    check_parity(idx, orc, reads, True, label=f"lane lookup {case} {kernel}")
    monkeypatch.setenv("ZG_SWEEP_CHUNK_POSITIONS", "9000")
    check_parity(idx, orc, reads, True, label=f"lane lookup {case} {kernel}, chunks")


@pytest.mark.parametrize("records", [True, False])
def test_sweep_as_one_graph_replays(records, monkeypatch):
    # the sweep of a batch captured as one HIP graph after its first (eager)
    # run and replayed: every replay, with reseeding and host work between,
    # equals the eager run and the oracle; several chunks per replay
    colors = make_colors(520, 560, 2500)
    reads = make_reads(521, colors, 150, lo=30, hi=2500)
    idx = build(colors, fingerprint_bits=8, hashes=4, slot_scale=0.6)
    idx.load(recover=True)
    orc = oracle_for(idx, True)
    buf, off = zindex_amd.pack_reads(reads)
    exp_totals, exp_counts = orc.query_reads(buf, off, threads=zr.cpu_count())
    monkeypatch.setenv("ZG_SWEEP_CHUNK_POSITIONS", "40000")
    idx.tuning = dict(SWEEP_BENCH, graph=1)
    with idx.batch(buf, off, records=records) as b:
        b.run()  # eager, then captured
        first = b.results()
        st = b.stats()
        assert st["sweep_launches"] > 1 and st["tuning"]["graph"] == 1, st
        for rep in range(4):
            if rep == 2:
                b.reseed()
            b.run()
            st = b.stats()
            assert st["tuning"]["graph"] == 1 and st["sweep_lookup_ms"] == -1, st
            got = b.results()
            say(f"graph replay {rep}: records={records}, {st['last_total_ms']:.2f} ms")
            assert np.array_equal(got[0], exp_totals)
            assert np.array_equal(got[2], exp_counts.sum(axis=0))
            assert np.array_equal(got[0], first[0]) and np.array_equal(got[2], first[2])
            if records:
                assert np.array_equal(got[1].astype(np.uint64), exp_counts)
    idx.tuning = dict(SWEEP_BENCH, graph=0)


def test_sweep_where_only_the_bit_planes_fit():
    # b = 16 at H = 8 over 300 colours: the staged lookup's window (8 x 64 rows
    # of a 512-byte slice) exceeds LDS, the bit-plane lookup's (256-byte
    # slices) fits, so the sweep runs over the bit planes only
    colors = make_colors(610, 300, 3000)
    reads = make_reads(611, colors, 150, lo=30, hi=2500)
    idx = build(colors, fingerprint_bits=16, hashes=8)
    idx.load()
    orc = oracle_for(idx)
    idx.tuning = SWEEP_BENCH
    st = assert_sweep_runs(idx, reads)
    assert st["sweep_plane_banks"] == idx.meta["n_banks"], st
    check_parity(idx, orc, reads, label="b16 H8, bit planes only")


@pytest.mark.parametrize("bits,hashes,stage", [(1, 4, "1"), (2, 3, "1"), (4, 4, "1"), (4, 8, "0"),
                                               (16, 2, "1"), (32, 3, "1"), (8, 8, "1")])
def test_bit_plane_lookup_across_widths(bits, hashes, stage, monkeypatch):
    # 600 colours: one slice per row at b = 1 and 4, two at b = 2 and 8,
    # three at b = 16 and more at b = 32; hit rows through LDS (stage 1,
    # whole row slices per store) or stored by each lane (stage 0)
    colors = make_colors(620 + bits + hashes, 600, 2000)
    reads = make_reads(621 + bits, colors, 120, lo=30, hi=1800)
    idx = build(colors, fingerprint_bits=bits, hashes=hashes)
    idx.load()
    orc = oracle_for(idx)
    idx.tuning = SWEEP_BENCH
    monkeypatch.setenv("ZG_SWEEP_ROW_STAGE", stage)
    st = assert_sweep_runs(idx, reads)
    assert st["sweep_plane_banks"] == idx.meta["n_banks"], st
    check_parity(idx, orc, reads, label=f"planes b{bits} H{hashes} stage{stage}")
