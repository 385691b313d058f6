"""GPU parity: the HIP query path (through the C ABI) against the CPU oracle,
bit-exact on per-read totals, per-read per-colour counts and aggregates."""
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np
import pytest

import zindex_ref as zr
from zindex_ref import GOLDEN, ROOT

zindex_amd = pytest.importorskip("zindex_amd")
from zindex_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_colors(seed, n_colors, length, rates=(0.01, 0.2)):
    rng = np.random.default_rng(seed)
    root = synth.random_genome(rng, length)
    colors = [root]
    for i in range(1, n_colors):
        colors.append(synth.substitute(rng, root, float(rng.uniform(*rates))))
    return [c.tobytes() for c in colors]


def make_reads(seed, colors, n, lo=20, hi=1500, error=0.02):
    rng = np.random.default_rng(seed)
    arrs = [np.frombuffer(c, np.uint8) for c in colors]
    lengths = rng.integers(lo, hi, n)
    buf, off, _ = synth.sample_reads(rng, arrs, n, lengths, error)
    reads = [buf[off[i]:off[i + 1]].tobytes() for i in range(n)]
    # edge cases: empty, shorter than k, N runs, lower case, all-N
    reads[0] = b""
    reads[1] = reads[1][:10]
    if len(reads[2]) > 100:
        r = bytearray(reads[2])
        r[40:45] = b"NNNNN"
        reads[2] = bytes(r)
    reads[3] = reads[3].lower()
    reads[4] = b"N" * 50
    return reads


def build(colors, **cfg):
    b = zindex_amd.Builder(**cfg)
    for i, seq in enumerate(colors):
        b.add_color(f"color{i}", [seq])
    return b.finish()


def oracle_for(idx, recover=True):
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "idx"
        idx.write(str(out))
        disk = zr.read_index_dir(out)
    return zr.OracleIndex(disk["manifest"], disk["banks"], disk["recovery"], recover)


def check_parity(idx, orc, reads, recover=True, label="", sweep=True):
    """Records and aggregate against the oracle with the index's tuning and,
    with sweep=True, again as the sorted sweep (where the index allows it)."""
    buf, off = zindex_amd.pack_reads(reads)
    say(f"parity {label}: {len(reads)} reads, {int(off[-1])} bases")
    exp_totals, exp_counts = orc.query_reads(buf, off, threads=zr.cpu_count())
    saved = idx.tuning
    tunings = [("", saved)]
    if sweep and not saved.get("sweep"):
        tunings.append((" [sweep]", dict(saved, sweep=1)))
    for tag, tuning in tunings:
        idx.tuning = tuning
        totals, counts, agg = idx.query_reads(buf, off, records=True, recover=recover)
        assert np.array_equal(totals, exp_totals), label + tag
        bad = np.argwhere(counts.astype(np.uint64) != exp_counts)
        assert len(bad) == 0, f"{label}{tag}: {len(bad)} mismatches, first {bad[:5].tolist()}"
        assert np.array_equal(agg, exp_counts.sum(axis=0)), label + tag
        t2, _, agg2 = idx.query_reads(buf, off, records=False, recover=recover)
        assert np.array_equal(t2, exp_totals) and np.array_equal(agg2, exp_counts.sum(axis=0)), \
            label + tag
    idx.tuning = saved
    return exp_counts


def assert_sweep_runs(idx, reads):
    """The batch runs as the sorted sweep (not the read-ordered lookup);
    returns its stats."""
    buf, off = zindex_amd.pack_reads(reads)
    saved = idx.tuning
    idx.tuning = dict(saved, sweep=1)
    with idx.batch(buf, off, records=True) as b:
        b.run()
        st = b.stats()
    idx.tuning = saved
    assert st["tuning"]["sweep"] == 1 and st["sweep_launches"] >= 1, st
    return st


@pytest.fixture(scope="module")
def base_case():
    colors = make_colors(1, 10, 40000)
    reads = make_reads(2, colors, 300)
    idx = build(colors, fingerprint_bits=8, hashes=4)
    idx.load()
    return colors, reads, idx, oracle_for(idx)


def test_kmers_b8_h4_parity(base_case):
    colors, reads, idx, orc = base_case
    counts = check_parity(idx, orc, reads, label="b8h4")
    assert counts.sum() > 0


@pytest.mark.parametrize("bits,hashes", [(1, 3), (2, 4), (4, 2), (16, 3), (32, 8), (8, 2),
                                         (4, 8), (16, 4)])
def test_width_and_hash_sweep(bits, hashes):
    colors = make_colors(10 + bits + hashes, 6, 25000)
    reads = make_reads(20 + bits, colors, 120)
    idx = build(colors, fingerprint_bits=bits, hashes=hashes)
    idx.load()
    check_parity(idx, oracle_for(idx), reads, label=f"b{bits}h{hashes}")


@pytest.mark.parametrize("group,bits", [(3, 8), (5, 4), (2, 16), (7, 1)])
def test_merged_groups(group, bits):
    colors = make_colors(30 + group, 9, 20000, rates=(0.0, 0.05))
    reads = make_reads(40 + group, colors, 150)
    idx = build(colors, fingerprint_bits=bits, hashes=4, merge_group_size=group)
    idx.load()
    check_parity(idx, oracle_for(idx), reads, label=f"merged g{group} b{bits}")


@pytest.mark.parametrize("bits,group", [(8, 1), (8, 3), (2, 1)])
def test_recovery_high_load(bits, group):
    colors = make_colors(50 + bits, 6, 30000)
    reads = make_reads(60, colors, 200)
    idx = build(colors, fingerprint_bits=bits, hashes=4, slot_scale=0.5, merge_group_size=group)
    assert idx.meta["abandoned_hashes"] > 0
    idx.load(recover=True)
    check_parity(idx, oracle_for(idx, True), reads, True, label=f"recovery b{bits} g{group}")
    idx2 = build(colors, fingerprint_bits=bits, hashes=4, slot_scale=0.5, merge_group_size=group)
    idx2.load(recover=False)
    check_parity(idx2, oracle_for(idx2, False), reads, False, label=f"no-recovery b{bits}")


def test_many_tiles_and_tile_grouping():
    # 1100 singleton colours at b=8: two 1024-column tiles per row
    rng = np.random.default_rng(70)
    root = synth.random_genome(rng, 3000)
    colors = [synth.substitute(rng, root, 0.02).tobytes() for _ in range(1100)]
    reads = make_reads(71, colors, 60, lo=40, hi=600)
    idx = build(colors, fingerprint_bits=8, hashes=4)
    idx.load()
    orc = oracle_for(idx)
    for tiles, prefetch, nt in ((1, 1, 0), (1, 4, 1), (2, 1, 1), (2, 4, 0), (2, 2, 1)):
        idx.tuning = {"span": 1024, "tiles_per_wave": tiles, "prefetch": prefetch,
                      "waves_per_cu": 0, "nontemporal": nt}
        check_parity(idx, orc, reads, label=f"tiles{tiles} d{prefetch} nt{nt}")


@pytest.mark.parametrize("bits,group,scale", [(1, 1, 1.0), (2, 3, 0.5), (4, 1, 0.5)])
def test_deep_prefetch_narrow_rows(bits, group, scale):
    colors = make_colors(75 + bits, 8, 20000)
    reads = make_reads(76 + bits, colors, 150)
    idx = build(colors, fingerprint_bits=bits, hashes=3, merge_group_size=group, slot_scale=scale)
    idx.load()
    orc = oracle_for(idx)
    for depth in (8, 16):
        for span in (64, 1024):
            idx.tuning = {"span": span, "tiles_per_wave": 1, "prefetch": depth,
                          "waves_per_cu": 0, "nontemporal": 0}
            check_parity(idx, orc, reads, label=f"b{bits} g{group} d{depth} span{span}")


@pytest.mark.parametrize("n_colors,bits,group,mode", [(37, 4, 3, 0), (13, 8, 1, 1), (70, 16, 2, 3),
                                                     (5, 32, 1, 0), (99, 2, 4, 2)])
def test_padded_row_pitch(n_colors, bits, group, mode):
    # row lengths that are not multiples of 16 bytes are padded in HBM
    colors = make_colors(190 + n_colors, n_colors, 6000)
    reads = _with_runs(make_reads(191 + n_colors, colors, 120, lo=20, hi=1500))
    cfg = dict(fingerprint_bits=bits, hashes=3, merge_group_size=group, slot_scale=0.7,
               index_mode=mode, k=31, minimizer_size=19, modimizer_sampling=3, findere_z=6)
    idx = build(colors, **cfg)
    rb = [(b["row_bits"] + 7) // 8 for b in idx.manifest()["banks"]]
    assert any(r % 16 for r in rb), rb
    idx.load()
    check_parity(idx, oracle_for(idx), reads, label=f"pitch c{n_colors} b{bits} g{group} m{mode}")


SWEEP = {"span": 1024, "tiles_per_wave": 1, "prefetch": 2, "waves_per_cu": 0,
         "nontemporal": 0, "sweep": 1}



@pytest.mark.parametrize("bits,group,scale,mode,hpc,pattern", [
    (8, 1, 1.0, 0, False, None), (8, 3, 0.5, 0, False, None), (4, 2, 0.5, 2, False, None),
    (16, 1, 1.0, 0, True, "###_##_#_##_#_###_#_##_#_##_###"), (1, 4, 0.6, 0, False, None)])
def test_sorted_sweep_parity(bits, group, scale, mode, hpc, pattern, monkeypatch):
    # seeds looked up in PositionHash segment order, hits summed per read
    colors = make_colors(200 + bits + group, 9, 20000)
    reads = _with_runs(make_reads(201 + bits, colors, 200, lo=20, hi=2500))
    idx = build(colors, fingerprint_bits=bits, hashes=4, merge_group_size=group,
                slot_scale=scale, index_mode=mode, modimizer_sampling=3,
                homopolymer_compression=int(hpc), seed_pattern=pattern or "")
    idx.load()
    orc = oracle_for(idx)
    idx.tuning = SWEEP
    assert_sweep_runs(idx, reads)
    check_parity(idx, orc, reads, label=f"sweep b{bits} g{group} m{mode}")
    # several chunks of whole reads
    monkeypatch.setenv("ZG_SWEEP_CHUNK_POSITIONS", "5000")
    check_parity(idx, orc, reads, label=f"sweep chunks b{bits} g{group} m{mode}")


def test_store_probe_needs_a_sweep_and_keeps_results(base_case):
    # zg_batch_store_probe replays the last chunk's hit-row stores: refused
    # before any sweep, a positive rate after one, and the batch's results
    # (per-read counts) unchanged by it
    colors, reads, idx, orc = base_case
    buf, off = zindex_amd.pack_reads(reads)
    saved = idx.tuning
    idx.tuning = dict(SWEEP)
    try:
        with idx.batch(buf, off, records=True) as b:
            with pytest.raises(zindex_amd._lib.ZindexError):
                b.store_probe()
            b.run()
            before = b.results()
            gbs = b.store_probe()
            after = b.results()
    finally:
        idx.tuning = saved
    say(f"store probe: {gbs:.1f} GB/s")
    assert gbs > 0
    assert all(np.array_equal(x, y) for x, y in zip(before, after))


def test_sweep_pieces_against_cpu():
    # the sweep's scans, stable radix passes by first segment, segment starts
    # and per-read row sums, each against its CPU restatement
    exe = ROOT / "tests" / "sweep_pieces"
    assert exe.exists(), "tests/sweep_pieces is built by __graft_entry__.build()"
    say("sweep pieces: scans, radix passes, segment starts, row sums")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all pieces match" in r.stdout


def test_minimizer_positions_and_counts():
    colors = make_colors(80, 5, 30000)
    reads = make_reads(81, colors, 120, lo=30, hi=3000, error=0.05)
    idx = build(colors, fingerprint_bits=16, hashes=3, index_mode=1, k=49, minimizer_size=31)
    idx.load()
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, pos = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp_pos, exp_keys = zr.scan_ocm(r, 49, 31, 69)
        a, b = int(key_off[i]), int(key_off[i + 1])
        assert pos[a:b].tolist() == exp_pos.tolist(), i
        assert keys[a:b].tolist() == exp_keys.tolist(), i
    assert_sweep_runs(idx, reads)
    check_parity(idx, oracle_for(idx), reads, label="ocm k49 m31 b16 h3")


@pytest.mark.parametrize("k,m", [(1, 1), (7, 7), (19, 12), (63, 21), (255, 31), (31, 4)])
def test_minimizer_parameter_edges(k, m):
    colors = make_colors(90 + k, 3, 8000)
    reads = make_reads(91 + m, colors, 40, lo=1, hi=800)
    idx = build(colors, fingerprint_bits=8, hashes=3, index_mode=1, k=k, minimizer_size=m)
    idx.load()
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, pos = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp_pos, exp_keys = zr.scan_ocm(r, k, m, 69)
        a, b = int(key_off[i]), int(key_off[i + 1])
        assert pos[a:b].tolist() == exp_pos.tolist(), (k, m, i)
        assert keys[a:b].tolist() == exp_keys.tolist(), (k, m, i)
    check_parity(idx, oracle_for(idx), reads, label=f"ocm k{k} m{m}")


@pytest.mark.parametrize("sampling,bits", [(1, 8), (5, 4), (16, 16)])
def test_modimizer_seeds_and_counts(sampling, bits):
    colors = make_colors(120 + sampling, 6, 30000)
    reads = make_reads(121 + sampling, colors, 150, lo=20, hi=2500)
    idx = build(colors, fingerprint_bits=bits, hashes=3, index_mode=2,
                modimizer_sampling=sampling)
    idx.load()
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, _ = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp = zr.read_features(r, 31, 69, mode=2, sampling=sampling)
        assert keys[int(key_off[i]):int(key_off[i + 1])].tolist() == exp.tolist(), i
    check_parity(idx, oracle_for(idx), reads, label=f"modimizers s{sampling} b{bits}")


K31_PATTERN = "###_##_#_##_#_###_#_##_#_##_###"
M19_PATTERN = "#####_#_###_#_#####"


def _with_runs(reads):
    """Long homopolymer runs, mixed case and N breaks inside some reads."""
    out = list(reads)
    for i in range(5, len(out), 7):
        r = bytearray(out[i])
        if len(r) > 120:
            r[30:50] = b"AAAAAAaaaaCCCCccccGG"
            r[70:75] = b"TtTtN"
            r[100:104] = b"gggg"
        out[i] = bytes(r)
    return out


@pytest.mark.parametrize("mode,hpc,pattern", [
    (0, True, None), (1, True, None), (2, True, None),
    (0, False, K31_PATTERN), (1, False, M19_PATTERN), (1, True, M19_PATTERN)])
def test_compression_and_spaced_seed_parity(mode, hpc, pattern):
    colors = make_colors(140 + mode, 6, 30000)
    reads = _with_runs(make_reads(141 + mode, colors, 160, lo=20, hi=2500))
    idx = build(colors, fingerprint_bits=8, hashes=3, index_mode=mode, k=31, minimizer_size=19,
                modimizer_sampling=3, homopolymer_compression=int(hpc),
                seed_pattern=pattern or "")
    idx.load()
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, _ = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp = zr.read_features(r, 31, 69, mode=mode, m=19, sampling=3, hpc=hpc, pattern=pattern)
        assert keys[int(key_off[i]):int(key_off[i + 1])].tolist() == exp.tolist(), i
    check_parity(idx, oracle_for(idx), reads, label=f"mode{mode} hpc{hpc} {pattern}")


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_compressed_index_answers_compressed_and_expanded_queries(mode):
    # tests/minimizer_schemes_cli.rs:102-177: an index of an expanded sequence
    # built with compression finds every seed of the compressed sequence, of its
    # reverse complement and of the expanded sequence itself
    rng = np.random.default_rng(150 + mode)
    raw = synth.random_genome(rng, 1800).tobytes()
    compressed = zr.homopolymer_compress(raw)
    expanded = b"".join(bytes([c]) * int(n) for c, n in
                        zip(compressed, rng.integers(1, 4, len(compressed))))
    rc = bytes(b"TGCA"[b"ACGT".index(c)] for c in reversed(compressed))
    idx = build([expanded], fingerprint_bits=8, hashes=3, index_mode=mode, k=31,
                minimizer_size=21, modimizer_sampling=16, homopolymer_compression=1)
    idx.load()
    expected = len(zr.read_features(compressed, 31, 69, mode=mode, m=21, sampling=16))
    assert expected > 0
    for query in (compressed, rc, expanded):
        buf, off = zindex_amd.pack_reads([query])
        totals, counts, agg = idx.query_reads(buf, off, records=True)
        assert int(totals[0]) == expected and int(counts[0][0]) == expected


@pytest.mark.parametrize("bits,group,k,z,hpc,scale", [
    (8, 1, 31, 10, False, 1.0), (16, 2, 41, 10, False, 1.0), (32, 1, 31, 10, False, 1.0),
    (4, 3, 31, 5, True, 1.0), (8, 2, 25, 0, False, 0.5), (1, 1, 90, 60, False, 1.0)])
def test_findere_parity(bits, group, k, z, hpc, scale):
    colors = make_colors(160 + bits + k, 7, 30000)
    reads = _with_runs(make_reads(161 + z, colors, 150, lo=20, hi=2500))
    idx = build(colors, fingerprint_bits=bits, hashes=3, index_mode=3, k=k, findere_z=z,
                merge_group_size=group, homopolymer_compression=int(hpc), slot_scale=scale)
    if scale < 1:
        assert idx.meta["abandoned_hashes"] > 0
    idx.load()
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, _ = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp = zr.scan_kmers(zr.homopolymer_compress(r) if hpc else r, k - z)
        assert keys[int(key_off[i]):int(key_off[i + 1])].tolist() == exp.tolist(), i
    label = f"findere b{bits} g{group} k{k} z{z} hpc{hpc} scale{scale}"
    assert check_parity(idx, oracle_for(idx), reads, label=label).sum() > 0
    if group == 1 and not hpc:
        idx.tuning = {"span": 256, "tiles_per_wave": 1, "prefetch": 1, "waves_per_cu": 0,
                      "nontemporal": 0}
        check_parity(idx, oracle_for(idx), reads, label=label + " span256")


def _revcomp(seq: bytes) -> bytes:
    comp = {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A")}
    return bytes(comp.get(ord(chr(c).upper()), c) for c in reversed(seq))


@pytest.mark.parametrize("mode,pattern,bounds,hpc,bits", [
    (0, None, (0, 0), False, 8), (0, K31_PATTERN, (10, 50), True, 16),
    (1, None, (0, 0), False, 8), (1, M19_PATTERN, (5, 40), False, 4),
    (0, None, (1, 1), False, 32)])
def test_randstrobe_parity(mode, pattern, bounds, hpc, bits):
    colors = make_colors(170 + mode + bits, 6, 30000)
    reads = _with_runs(make_reads(171 + bits, colors, 150, lo=20, hi=2500))
    idx = build(colors, fingerprint_bits=bits, hashes=3, index_mode=mode, k=31, minimizer_size=19,
                randstrobemers=1, strobe_min=bounds[0], strobe_max=bounds[1],
                homopolymer_compression=int(hpc), seed_pattern=pattern or "")
    idx.load()
    strobes = (idx.manifest()["randstrobemers"]["min_distance"],
               idx.manifest()["randstrobemers"]["max_distance"])
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, _ = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp = zr.read_features(r, 31, 69, mode=mode, m=19, hpc=hpc, pattern=pattern,
                               strobes=strobes)
        assert keys[int(key_off[i]):int(key_off[i + 1])].tolist() == exp.tolist(), i
    label = f"randstrobes mode{mode} {pattern} {strobes} hpc{hpc} b{bits}"
    assert_sweep_runs(idx, reads)
    assert check_parity(idx, oracle_for(idx), reads, label=label).sum() > 0
    # the count is the larger of the two orientations, so a read and its
    # reverse complement score the same
    pair = [reads[10], _revcomp(reads[10])]
    t, c, _ = idx.query_reads(*zindex_amd.pack_reads(pair), records=True)
    assert t[0] == t[1] and np.array_equal(c[0], c[1]) and c[0].sum() > 0


@pytest.mark.parametrize("k,mode,hpc", [(33, 0, False), (64, 0, False), (65, 0, True),
                                         (255, 0, False), (511, 0, False), (45, 2, False)])
def test_long_kmer_parity(k, mode, hpc):
    # tests/long_kmer_cli.rs: k-mers longer than 32 hash both orientations
    colors = make_colors(180 + k, 5, 20000)
    reads = make_reads(181 + k, colors, 120, lo=20, hi=2500, error=0.002)
    idx = build(colors, fingerprint_bits=8, hashes=4, k=k, index_mode=mode, modimizer_sampling=4,
                homopolymer_compression=int(hpc))
    idx.load()
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, pos = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp = zr.read_features(r, k, 69, mode=mode, sampling=4, hpc=hpc)
        assert keys[int(key_off[i]):int(key_off[i + 1])].tolist() == exp.tolist(), i
    label = f"long k{k} mode{mode} hpc{hpc}"
    assert_sweep_runs(idx, reads)
    assert check_parity(idx, oracle_for(idx), reads, label=label).sum() > 0
    # a source and its reverse complement match every one of their k-mers
    src = colors[0][1000:4000]
    t, c, _ = idx.query_reads(*zindex_amd.pack_reads([src, _revcomp(src), b"ACGT"]), records=True)
    assert t[0] == t[1] and c[0][0] == t[0] and c[1][0] == t[1] and t[2] == 0


def test_kmer_seeds_match_oracle_scan(base_case):
    colors, reads, idx, orc = base_case
    buf, off = zindex_amd.pack_reads(reads)
    key_off, keys, pos = idx.extract_seeds(buf, off)
    for i, r in enumerate(reads):
        exp = zr.scan_kmers(r, 31)
        a, b = int(key_off[i]), int(key_off[i + 1])
        assert keys[a:b].tolist() == exp.tolist(), i


def test_plain_path_graph_replays_and_autotune_agree(base_case):
    colors, reads, idx, orc = base_case
    buf, off = zindex_amd.pack_reads(reads)
    exp_totals, exp_counts = orc.query_reads(buf, off)
    with idx.batch(buf, off, records=True) as b:
        for rep in range(3):  # graph captured on the first run, replayed after
            b.run()
            totals, counts, agg = b.results()
            assert np.array_equal(counts.astype(np.uint64), exp_counts), rep
            _ = np.zeros(10).sum()  # host work between replays
        b.run_plain()
        totals, counts, agg = b.results()
        assert np.array_equal(counts.astype(np.uint64), exp_counts), "plain"
        chosen = b.autotune()
        b.run()
        totals, counts, agg = b.results()
        assert np.array_equal(counts.astype(np.uint64), exp_counts), f"autotuned {chosen}"
        st = b.stats()
        assert st["seeds"] == int(exp_totals.sum()) and st["last_lookup_ms"] > 0
        # the batch alternates between the sorted sweep and the read-ordered lookup
        # (whose runs free the sweep's buffers)
        tuned = idx.tuning
        for sweep in (1, 0, 1, 0):
            idx.tuning = {**tuned, "sweep": sweep}
            b.run()
            _, counts, _ = b.results()
            assert np.array_equal(counts.astype(np.uint64), exp_counts), f"sweep {sweep}"
        idx.tuning = tuned


def test_aggregate_batch_and_stats(base_case):
    colors, reads, idx, orc = base_case
    buf, off = zindex_amd.pack_reads(reads)
    _, exp_counts = orc.query_reads(buf, off)
    with idx.batch(buf, off, records=False) as b:
        b.run()
        b.run()
        totals, _, agg = b.results()
        assert np.array_equal(agg, exp_counts.sum(axis=0))


@pytest.fixture(scope="module")
def bacteria_index(tmp_path_factory):
    d = tmp_path_factory.mktemp("bacteria")
    out = d / "idx"
    say("building the 10-genome bacteria index on the host")
    zindex_amd.build(str(GOLDEN / "bacteria" / "samples.fof"), str(out), k=31)
    return out


def _run_cli(args):
    cli = ROOT / "zindex_amd" / "zorindex"
    r = subprocess.run([str(cli)] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_bacteria_readme_and_demo_known_answers(bacteria_index):
    v = zr.vectors()["readme_bacteria"]
    orc = zr.OracleIndex.from_dir(bacteria_index)
    colors = [(c["name"], c["unique_kmers"]) for c in orc.manifest["colors"]]
    # aggregate query.fa: byte-identical to the oracle's restated output
    out = _run_cli(["query", "-i", str(bacteria_index), "-q", str(GOLDEN / "bacteria" / "query.fa")])
    recs = zr.read_fasta_records(GOLDEN / "bacteria" / "query.fa")
    buf, off = zindex_amd.pack_reads([s for _, s in recs])
    totals, counts = orc.query_reads(buf, off)
    assert out == zr.aggregate_text(colors, counts.sum(axis=0), int(totals.sum()), 0.0)
    src = [c[0] for c in colors].index(v["query_fa_source"])
    assert f"{v['query_fa_source']}\t{v['query_fa_matches']}\t1.00000000\t" in out
    # records with --min-ratio 0.8 (demo.sh:30-44): every positive matches its source fully
    qpath = GOLDEN / "bacteria" / "queries.fa"
    out = _run_cli(["query", "-i", str(bacteria_index), "-q", str(qpath), "--records",
                    "--min-ratio", "0.8"])
    recs = zr.read_fasta_records(qpath)
    buf, off = zindex_amd.pack_reads([s for _, s in recs])
    totals, counts = orc.query_reads(buf, off)
    names = [n for n, _ in recs]
    assert out == zr.records_text(names, totals, counts, colors, 0.8, "tsv")
    truth = [l.split("\t") for l in (GOLDEN / "bacteria" / "source-truth.tsv").read_text().splitlines()[1:]]
    for qid, doc, *_ in truth:
        i = names.index(qid)
        if doc == "-":  # the random negative control has no source
            continue
        assert int(counts[i, int(doc)]) == int(totals[i]) > 0, qid
    for first in v["records_first_two"]:
        assert "\t".join([first[0], str(first[1]), colors[first[1]][0], str(first[2]),
                          str(first[3]), "1.00000000"]) in out
    # compact format and ratio 0
    out = _run_cli(["query", "-i", str(bacteria_index), "-q", str(qpath), "--records",
                    "-F", "compact"])
    assert out == zr.records_text(names, totals, counts, colors, 0.0, "compact")


@pytest.mark.parametrize("codec", ["plain", "gzip", "zstd", "xz", "bzip2", "fastq"])
def test_streamed_query_file_in_record_batches(bacteria_index, tmp_path, codec):
    """The query file is read and decompressed chunk by chunk and parsed in
    batches of -B records on a thread of its own (records.rs:58-148): every
    batch size and codec (concatenated members, FASTQ) prints exactly the
    oracle's text; a truncated stream fails."""
    import bz2
    import gzip
    import lzma
    qpath = GOLDEN / "bacteria" / "queries.fa"
    recs = zr.read_fasta_records(qpath)
    raw = qpath.read_bytes()
    if codec == "fastq":
        raw = b"".join(b"@%s extra words\n%s\n+\n%s\n" % (n.encode(), s, b"I" * len(s))
                       for n, s in recs)
    half = len(raw) // 2
    enc = {"plain": lambda b: b, "fastq": lambda b: b,
           "gzip": lambda b: gzip.compress(b[:half]) + gzip.compress(b[half:]),
           "zstd": lambda b: zr.zstd_compress(b),
           "xz": lambda b: lzma.compress(b[:half]) + lzma.compress(b[half:]),
           "bzip2": lambda b: bz2.compress(b[:half]) + bz2.compress(b[half:])}[codec]
    path = tmp_path / "q.x"
    path.write_bytes(enc(raw))
    orc = zr.OracleIndex.from_dir(bacteria_index)
    colors = [(c["name"], c["unique_kmers"]) for c in orc.manifest["colors"]]
    buf, off = zindex_amd.pack_reads([s for _, s in recs])
    totals, counts = orc.query_reads(buf, off)
    names = [n for n, _ in recs]
    want = zr.records_text(names, totals, counts, colors, 0.0, "compact")
    for batch in (1, 7, 4096):
        out = _run_cli(["query", "-i", str(bacteria_index), "-q", str(path), "--records",
                        "-F", "compact", "-B", str(batch)])
        assert out == want, (codec, batch)
    out = _run_cli(["query", "-i", str(bacteria_index), "-q", str(path)])
    assert out == zr.aggregate_text(colors, counts.sum(axis=0), int(totals.sum()), 0.0)
    if codec in ("gzip", "xz", "bzip2", "zstd"):
        bad = tmp_path / "bad.x"
        bad.write_bytes(path.read_bytes()[:-12])
        cli = ROOT / "zindex_amd" / "zorindex"
        r = subprocess.run([str(cli), "query", "-i", str(bacteria_index), "-q", str(bad),
                            "--records", "-B", "3"], capture_output=True, text=True, timeout=300)
        assert r.returncode != 0 and codec in r.stderr, r.stderr


def test_bacteria_minimizer_demo(tmp_path):
    out = tmp_path / "min_idx"
    zindex_amd.build(str(GOLDEN / "bacteria" / "samples.fof"), str(out), index_mode=1, k=49,
                     minimizer_size=31)
    orc = zr.OracleIndex.from_dir(out)
    qpath = GOLDEN / "bacteria" / "queries.fa"
    recs = zr.read_fasta_records(qpath)
    names = [n for n, _ in recs]
    buf, off = zindex_amd.pack_reads([s for _, s in recs])
    totals, counts = orc.query_reads(buf, off)
    text = _run_cli(["query", "-i", str(out), "-q", str(qpath), "--records", "--min-ratio", "0.8"])
    colors = [(c["name"], c["unique_kmers"]) for c in orc.manifest["colors"]]
    assert text == zr.records_text(names, totals, counts, colors, 0.8, "tsv")
    truth = [l.split("\t") for l in (GOLDEN / "bacteria" / "source-truth.tsv").read_text().splitlines()[1:]]
    for qid, doc, *_ in truth:
        i = names.index(qid)
        if doc == "-":  # the random negative control has no source
            continue
        assert int(counts[i, int(doc)]) == int(totals[i]) > 0, qid


def test_hbm_budget_rejects_oversized_index(base_case):
    colors, reads, idx, orc = base_case
    small = build(colors[:2], fingerprint_bits=8, hashes=4)
    with pytest.raises(zindex_amd.ZindexError) as e:
        small.load(hbm_budget=1024)
    assert e.value.code == -7


@pytest.mark.parametrize("group,cap,device", [(1, 1 << 15, 0), (3, 1 << 15, 0), (1, 1 << 14, 1)])
def test_hash_partitioned_index(group, cap, device):
    """Banks of an oversized size class split into hash partitions
    (storage.rs:1525-1641): each key is looked up in the partition partition_id
    routes it to (storage.rs:99-101, 2717-2745)."""
    colors = make_colors(110 + group, 8, 30000)
    reads = make_reads(111, colors, 200)
    idx = build(colors, fingerprint_bits=8, hashes=4, merge_group_size=group, slot_scale=0.8,
                max_subindex_size=cap, device=device)
    man = idx.manifest()
    assert man["version"] == 9 and idx.meta["n_banks"] > 1
    assert all(b.get("routing") for b in man["banks"])
    idx.load()
    assert_sweep_runs(idx, reads)
    check_parity(idx, oracle_for(idx), reads, label=f"partitioned g{group} cap {cap}")


@pytest.mark.parametrize("bits,group,reverse", [(8, 3, True), (4, 1, False), (16, 2, True)])
def test_legacy_interleaved_banks(tmp_path, bits, group, reverse):
    """Legacy interleaved rows (fingerprint_first = false, storage.rs:2864-2914),
    groups optionally out of width order: rewritten to the device plane layout
    on upload, answers equal the oracle's on the legacy files."""
    colors = make_colors(120 + bits, 7, 25000, rates=(0.0, 0.05))
    reads = make_reads(121 + bits, colors, 150)
    idx = build(colors, fingerprint_bits=bits, hashes=4, merge_group_size=group, slot_scale=0.8)
    plane, legacy = tmp_path / "plane", tmp_path / "legacy"
    idx.write(str(plane))
    zr.to_legacy_layout(plane, legacy,
                        (lambda gs: range(len(gs) - 1, -1, -1)) if reverse else None)
    disk = zr.read_index_dir(legacy)
    orc = zr.OracleIndex(disk["manifest"], disk["banks"], disk["recovery"])
    with zindex_amd.Index.open(str(legacy)) as li:
        li.load()
        check_parity(li, orc, reads, label=f"legacy b{bits} g{group}")


def test_legacy_segment_geometry():
    """No fixed segment length (manifest version 7): PositionHash's legacy
    geometry, length 2^floor(log2(clamp(slots/H, 1, 4096))) (storage.rs:3452-3481)."""
    colors = make_colors(130, 6, 30000)
    reads = make_reads(131, colors, 150)
    idx = build(colors, fingerprint_bits=8, hashes=3, segment_length=0, slot_scale=0.8)
    man = idx.manifest()
    assert man["version"] == 7 and man["segment_length"] is None
    idx.load()
    check_parity(idx, oracle_for(idx), reads, label="legacy geometry")


@pytest.fixture(scope="module")
def disk_index(tmp_path_factory):
    """An index on disk whose banks span several 4 MiB frames and 64 MiB upload
    chunks (80 colours x 1 Mbp, one 80 MiB bank), plus its reads."""
    colors = make_colors(140, 80, 1_000_000, rates=(0.005, 0.05))
    reads = make_reads(141, colors, 300)
    idx = build(colors, fingerprint_bits=8, hashes=4)
    out = tmp_path_factory.mktemp("disk") / "idx"
    idx.write(str(out))
    idx.load()
    expect = idx.query_reads(reads, records=True)
    return out, reads, expect


def test_streamed_load_from_disk(disk_index):
    """Bank files decode frame by frame into page-locked chunks copied to HBM
    (bank_io.rs:235-291); answers equal those of the index built in memory."""
    out, reads, expect = disk_index
    with zindex_amd.Index.open(str(out)) as z:
        z.load()
        got = z.query_reads(reads, records=True)
        for g, e in zip(got, expect):
            assert np.array_equal(g, e)


def test_subindex_mode_streams_banks_beyond_budget(tmp_path):
    """ZG_MEMORY_SUBINDEX: banks that do not fit the HBM budget stream in per
    query (QueryMemory::Subindex, mod.rs:52-59); results equal all-resident."""
    colors = make_colors(150, 10, 200_000)
    reads = make_reads(151, colors, 200)
    idx = build(colors, fingerprint_bits=8, hashes=4, max_subindex_size=1 << 19)
    man = idx.manifest()
    assert idx.meta["n_banks"] >= 4
    out = tmp_path / "idx"
    idx.write(str(out))
    bank = man["banks"][0]["slots"] * ((man["banks"][0]["row_bits"] + 7) // 8)
    with zindex_amd.Index.open(str(out)) as full:
        full.load()
        expect = full.query_reads(reads, records=True)
        rec_bytes = full.meta["abandoned_hashes"] * 28 + full.meta["abandoned_origins"] * 8
    # room for the recovery map, the streaming buffer and one resident bank
    budget = rec_bytes + 2 * (bank + 4096) + 1024
    with zindex_amd.Index.open(str(out)) as z:
        with pytest.raises(zindex_amd.ZindexError) as e:
            z.load(hbm_budget=budget, memory="all")
        assert e.value.code == -7
        with pytest.raises(zindex_amd.ZindexError) as e:
            z.load(hbm_budget=bank, memory="subindex")
        assert e.value.code == -7
        z.load(hbm_budget=budget, memory="subindex")
        for _ in range(2):  # the second query starts from the bank left in HBM
            got = z.query_reads(reads, records=True)
            for g, x in zip(got, expect):
                assert np.array_equal(g, x)
        with z.batch(*zindex_amd.pack_reads(reads), records=True) as b:
            b.run()
            b.run()
            assert np.array_equal(b.results()[1], expect[1])
            b.run_plain()
            assert np.array_equal(b.results()[1], expect[1])
        # a server over the same index: each batch streams the banks beyond
        # the budget (one batch at a time); every request equals its rows of
        # the all-resident answer
        cuts = [0, 1, 37, 38, 120, 200]
        with zindex_amd.Server(z, slots=2, max_requests=2) as srv:
            tickets = [srv.submit(reads[a:b]) for a, b in zip(cuts[:-1], cuts[1:])]
            for (a, b), t in zip(zip(cuts[:-1], cuts[1:]), tickets):
                tot, cnt, agg, st = srv.wait(t)
                assert st["status"] == 0
                assert np.array_equal(tot, expect[0][a:b])
                assert np.array_equal(cnt, expect[1][a:b])
                assert np.array_equal(agg, expect[1][a:b].astype(np.uint64).sum(axis=0))
            assert srv.stats()["batches"] >= 3
        with pytest.raises(zindex_amd.ZindexError) as e:
            zindex_amd.Server(z, hbm_budget=budget)  # no room left for a slot
        assert e.value.code == -7
    # the CLI's -M subindex under the same budget prints what -M all prints
    q = tmp_path / "q.fa"
    q.write_bytes(b"".join(b">r%d\n%s\n" % (i, r) for i, r in enumerate(reads) if r))
    cli = str(ROOT / "zindex_amd" / "zorindex")
    for extra in ([], ["-r"]):
        a = subprocess.run([cli, "query", "-i", str(out), "-q", str(q), "-M", "all"] + extra,
                           capture_output=True, check=True).stdout
        b = subprocess.run([cli, "query", "-i", str(out), "-q", str(q), "-M", "subindex",
                            "--hbm-budget", str(budget)] + extra,
                           capture_output=True, check=True).stdout
        assert a == b and len(a) > 0


def test_fused_index_components(tmp_path):
    """Version-10 fused indexes (fusion.rs:233-480): each input a component,
    dataset ids offset, recovery maps merged. GPU counts equal the oracle's
    and each input's own GPU counts side by side."""
    colors = make_colors(160, 9, 30000)
    reads = make_reads(161, colors, 200)
    parts = [colors[:3], colors[3:6], colors[6:]]
    dirs, alone = [], []
    for i, cols in enumerate(parts):
        idx = build(cols, fingerprint_bits=8, hashes=4, slot_scale=0.75,
                    merge_group_size=3 if i == 2 else 1)
        d = tmp_path / f"in{i}"
        idx.write(str(d))
        dirs.append(d)
        idx.load()
        alone.append(idx.query_reads(reads, records=True))
    out = tmp_path / "fused"
    zindex_amd.fuse_many_indexes(dirs, out)
    disk = zr.read_index_dir(out)
    orc = zr.OracleIndex(disk["manifest"], disk["banks"], disk["recovery"])
    with zindex_amd.Index.open(str(out)) as z:
        assert z.meta["version"] == 10
        z.load()
        assert_sweep_runs(z, reads)
        check_parity(z, orc, reads, label="fused")
        t, c, _ = z.query_reads(reads, records=True)
        assert np.array_equal(c, np.concatenate([a[1] for a in alone], axis=1))
        assert all(np.array_equal(t, a[0]) for a in alone)


@pytest.mark.parametrize("group", [1, 3])
def test_version1_index(tmp_path, group):
    """Manifest version 1 (storage.rs:2555-2612, 2665-2674): interleaved banks
    with the groups' fallback keys after the payload; their recovery map is
    built at load. GPU counts equal the oracle's on the v1 files."""
    colors = make_colors(170 + group, 7, 25000)
    reads = make_reads(171, colors, 200)
    idx = build(colors, fingerprint_bits=8, hashes=4, merge_group_size=group,
                segment_length=0, slot_scale=0.9, drop_abandoned=1)
    plane = tmp_path / "plane"
    idx.write(str(plane))
    man = idx.manifest()
    rng = np.random.default_rng(172)
    keys = np.unique(np.concatenate([zr.scan_kmers(r, 31) for r in reads if len(r) >= 31]))
    fallback = {}
    for key in rng.choice(keys, 400, replace=False):
        bi = int(rng.integers(len(man["banks"])))
        gi = int(rng.integers(len(man["banks"][bi]["groups"])))
        n = len(man["banks"][bi]["groups"][gi]["members"])
        fallback.setdefault((bi, gi), []).append((int(key), int(rng.integers(1, 1 << n))))
    v1 = tmp_path / "v1"
    zr.to_v1(plane, v1, fallback)
    disk = zr.read_index_dir(v1)
    orc = zr.OracleIndex(disk["manifest"], disk["banks"], disk["recovery"])
    with zindex_amd.Index.open(str(v1)) as z:
        z.load()
        assert z.meta["abandoned_hashes"] == 0  # v1 keeps no separate map
        check_parity(z, orc, reads, label=f"version 1 g{group}")
