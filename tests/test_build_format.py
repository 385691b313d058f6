"""CPU tests of the product's host side: the C ABI exports, index
construction in the reference format, manifest validation and info. None of
these launch a kernel."""
import ctypes as C
import json
import os
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import zindex_ref as zr
from zindex_ref import ROOT, GOLDEN, oracle, vectors

zindex_amd = pytest.importorskip("zindex_amd")
from zindex_amd import synth  # noqa: E402


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "zindex_gpu.h").read_text()
    names = set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(zg_\w+)\s*\(", header, re.M))
    assert len(names) >= 25
    lib = C.CDLL(str(ROOT / "zindex_amd" / "libzindex_amd.so"))
    for name in sorted(names):
        assert hasattr(lib, name), name
    assert b"gfx950" in zindex_amd._lib.lib().zg_version()


def test_first_hash_inverse(tmp_path):
    """The sorted lookup lists each seed as its first PositionHash word
    h = mixsplit(key, seed) (storage.rs:3458-3465) and recovers the key as
    unsplitmix64(h) - seed (zg_hash.h). splitmix64 is a bijection of 64-bit
    words, so the round trip holds both ways for any key and seed, and the
    multipliers' inverses are exact modulo 2^64."""
    from zindex_amd.buildlib import HIPCC
    src = tmp_path / "inverse.cpp"
    src.write_text(r"""
#include "zg_hash.h"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 rng(11);
  long bad = 0;
  for (int i = 0; i < 200000; i++) {
    const uint64_t k = rng(), s = rng();
    bad += zg::unsplitmix64(zg::mixsplit(k, s)) - s != k;
    bad += zg::mixsplit(zg::unsplitmix64(k) - s, s) != k;
  }
  for (uint64_t k : {0ULL, 1ULL, ~0ULL}) bad += zg::unsplitmix64(zg::splitmix64(k)) != k;
  bad += zg::inverse_odd(0x94D049BB133111EBULL) * 0x94D049BB133111EBULL != 1;
  bad += zg::inverse_odd(0xBF58476D1CE4E5B9ULL) * 0xBF58476D1CE4E5B9ULL != 1;
  printf("%ld\n", bad);
  return bad != 0;
}
""")
    exe = tmp_path / "inverse"
    subprocess.run([HIPCC, "-O1", "-I", str(ROOT / "zindex_amd" / "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr


def _random_genome(rng, n):
    return synth.random_genome(rng, n).tobytes()


def _mutate(rng, seq, rate):
    return synth.substitute(rng, np.frombuffer(seq, np.uint8), rate).tobytes()


@pytest.fixture(scope="module")
def small_colors():
    rng = np.random.default_rng(7)
    root = _random_genome(rng, 30000)
    colors = [root] + [_mutate(rng, root, r) for r in (0.01, 0.05, 0.1, 0.2)]
    colors.append(_random_genome(rng, 20000) + b"N" + _random_genome(rng, 5000))
    return colors


def _build(colors, **cfg):
    b = zindex_amd.Builder(**cfg)
    for i, seq in enumerate(colors):
        b.add_color(f"c{i}", [seq])
    return b.finish()


@pytest.mark.parametrize("bits,hashes,group", [(8, 4, 1), (1, 3, 1), (2, 2, 1), (4, 4, 2),
                                               (16, 3, 1), (32, 8, 3), (8, 4, 3)])
def test_build_round_trip_no_false_negatives(tmp_path, small_colors, bits, hashes, group):
    idx = _build(small_colors, fingerprint_bits=bits, hashes=hashes, merge_group_size=group,
                 slot_scale=1.0)
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert man["fingerprint_bits"] == bits and man["hashes"] == hashes
    assert len(man["colors"]) == len(small_colors)
    for i, payload in enumerate(disk["banks"]):
        assert np.array_equal(payload, idx.bank_payload(i))
    # manifest passes the product's own validation when reopened
    with zindex_amd.Index.open(str(out)) as reopened:
        assert reopened.meta["n_colors"] == len(small_colors)
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(small_colors):
        keys = np.unique(zr.scan_kmers(seq, 31))
        assert man["colors"][c]["unique_kmers"] == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:3000]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))


@pytest.mark.parametrize("sampling", [1, 5])
def test_modimizer_build_inserts_sampled_kmers(tmp_path, small_colors, sampling):
    idx = _build(small_colors, fingerprint_bits=16, hashes=3, index_mode="modimizers",
                 modimizer_sampling=sampling)
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert man["index_mode"] == 2 and man["modimizer_sampling"] == sampling
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(small_colors):
        every = np.unique(zr.scan_kmers(seq, 31))
        keys = np.unique(zr.read_features(seq, 31, man["seed"], mode=2, sampling=sampling))
        if sampling == 1:
            assert np.array_equal(keys, every)
        assert man["colors"][c]["unique_kmers"] == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:2000]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))


K31_PATTERN = "###_##_#_##_#_###_#_##_#_##_###"
M19_PATTERN = "#####_#_###_#_#####"


@pytest.mark.parametrize("mode,hpc,pattern", [
    (0, True, None), (1, True, None), (2, True, None),
    (0, False, K31_PATTERN), (1, False, M19_PATTERN), (0, True, K31_PATTERN)])
def test_build_with_compression_and_spaced_seeds(tmp_path, small_colors, mode, hpc, pattern):
    cfg = dict(fingerprint_bits=16, hashes=3, index_mode=mode, homopolymer_compression=int(hpc),
               modimizer_sampling=3)
    if mode == 1:
        cfg.update(k=31, minimizer_size=19)
    if pattern:
        cfg["seed_pattern"] = pattern
    idx = _build(small_colors, **cfg)
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert man["homopolymer_compression"] is hpc and man.get("seed_pattern") == pattern
    with zindex_amd.Index.open(str(out)) as reopened:
        assert reopened.meta["n_colors"] == len(small_colors)
    info = tmp_path / "info.txt"
    fd = os.open(info, os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
    try:
        zindex_amd.info_command(str(out), out_fd=fd)
    finally:
        os.close(fd)
    text = info.read_text()
    # tests/minimizer_schemes_cli.rs:157, src/anchors.rs:112-126
    assert f"homopolymer_compression\t{str(hpc).lower()}\n" in text
    assert f"seed_pattern\t{pattern or 'none'}\n" in text and "randstrobemers\tfalse\n" in text
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(small_colors):
        keys = np.unique(zr.read_features(seq, man["k"], man["seed"], mode=mode,
                                          m=man["minimizer_size"], sampling=3, hpc=hpc,
                                          pattern=pattern))
        assert man["colors"][c]["unique_kmers"] == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:1500]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))


def _findere_smers(seq, k, z):
    """Canonical (k - z)-mers of the runs of at least k valid bases
    (src/lib.rs:632-640)."""
    out = []
    for run in re.split(rb"[^ACGTacgt]+", seq):
        if len(run) >= k:
            out += zr.scan_kmers(run, k - z).tolist()
    return np.unique(np.array(out, np.uint64))


@pytest.mark.parametrize("k,z,group", [(31, 10, 1), (41, 10, 2), (25, 0, 1)])
def test_findere_build_indexes_smers_of_long_runs(tmp_path, small_colors, k, z, group):
    idx = _build(small_colors, fingerprint_bits=8, hashes=4, index_mode="findere", k=k,
                 findere_z=z, merge_group_size=group)
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert man["index_mode"] == 3 and man["findere_z"] == z and man["k"] == k
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(small_colors):
        keys = _findere_smers(seq, k, z)
        assert man["colors"][c]["unique_kmers"] == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:1500]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))
    info = tmp_path / "info.txt"
    fd = os.open(info, os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
    try:
        zindex_amd.info_command(str(out), out_fd=fd)
    finally:
        os.close(fd)
    text = info.read_text()
    assert f"index_mode\tfindere\n" in text
    assert f"findere_z\t{z}\nindexed_k\t{k - z}\n" in text  # storage.rs:3982-3991


@pytest.mark.parametrize("k,mode", [(33, 0), (49, 0), (64, 0), (65, 0), (255, 0), (45, 2)])
def test_long_kmer_build(tmp_path, small_colors, k, mode):
    idx = _build(small_colors, fingerprint_bits=8, hashes=4, k=k, index_mode=mode,
                 modimizer_sampling=4)
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert man["k"] == k
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(small_colors):
        keys = np.unique(zr.read_features(seq, k, man["seed"], mode=mode, sampling=4))
        assert man["colors"][c]["unique_kmers"] == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:1500]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))


@pytest.mark.parametrize("mode,pattern,bounds,hpc", [
    (0, None, (0, 0), False), (0, K31_PATTERN, (10, 50), True), (1, None, (0, 0), False),
    (1, M19_PATTERN, (5, 40), False)])
def test_randstrobe_build(tmp_path, small_colors, mode, pattern, bounds, hpc):
    cfg = dict(fingerprint_bits=16, hashes=3, index_mode=mode, randstrobemers=1,
               strobe_min=bounds[0], strobe_max=bounds[1], homopolymer_compression=int(hpc))
    if mode == 1:
        cfg.update(k=31, minimizer_size=19)
    if pattern:
        cfg["seed_pattern"] = pattern
    idx = _build(small_colors, **cfg)
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    length = 19 if mode == 1 else 31
    lo, hi = bounds if bounds[0] else (length, 3 * length)  # src/bin/zorindex.rs:469-485
    assert man["randstrobemers"] == {"min_distance": lo, "max_distance": hi}
    assert man["randstrobe_scheme"] == "strobealign-popcount8-symmetric64-downstream-v1"
    with zindex_amd.Index.open(str(out)) as reopened:
        assert reopened.meta["n_colors"] == len(small_colors)
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(small_colors):
        keys = np.unique(zr.read_features(seq, man["k"], man["seed"], mode=mode, m=19,
                                          hpc=hpc, pattern=pattern, strobes=(lo, hi)))
        assert man["colors"][c]["unique_kmers"] == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:1500]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))


@pytest.mark.parametrize("cfg,msg", [
    (dict(randstrobemers=1, strobe_min=10, strobe_max=5), "must not exceed"),
    (dict(index_mode=3, k=31, findere_z=31), "smaller than query k"),
    (dict(index_mode=3, k=60, findere_z=10), "longer than 32"),
    (dict(seed_pattern="##_"), "symmetric"),
    (dict(seed_pattern="#x#"), "only"),
    (dict(seed_pattern="___"), "at least one"),
    (dict(seed_pattern="#" * 21), "must match k"),
    (dict(index_mode=1, k=31, minimizer_size=21, seed_pattern="#" * 31), "minimizer size m"),
    (dict(index_mode=2, seed_pattern="#" * 31), "require --index-mode"),
])
def test_seed_option_validation(small_colors, cfg, msg):
    with pytest.raises(Exception, match=msg):
        _build(small_colors[:1], **cfg)


def test_compressed_inputs_build_the_same_index(tmp_path, small_colors):
    """src/lib.rs:560-571: inputs may be gzip, zstd, xz or bzip2 (here also as
    concatenated members); they index exactly as the plain files do."""
    import bz2
    import gzip
    import lzma
    cli = ROOT / "zindex_amd" / "zorindex"
    fa = [b">r%d\n%s\n" % (i, s[:20000]) for i, s in enumerate(small_colors[:4])]
    enc = [lambda b: gzip.compress(b[:500]) + gzip.compress(b[500:]),
           lambda b: zr.zstd_compress(b),
           lambda b: lzma.compress(b[:700]) + lzma.compress(b[700:]),
           lambda b: bz2.compress(b[:900]) + bz2.compress(b[900:])]
    plain, packed = [], []
    for i, (data, f) in enumerate(zip(fa, enc)):
        (tmp_path / f"p{i}.fa").write_bytes(data)
        (tmp_path / f"z{i}.fa.x").write_bytes(f(data))
        plain.append(f"c{i}\t{tmp_path / f'p{i}.fa'}\n")
        packed.append(f"c{i}\t{tmp_path / f'z{i}.fa.x'}\n")
    (tmp_path / "plain.fof").write_text("".join(plain))
    (tmp_path / "packed.fof").write_text("".join(packed))
    for fof, out in (("plain.fof", "ip"), ("packed.fof", "iz")):
        r = subprocess.run([str(cli), "build", "-f", str(tmp_path / fof), "-o", str(tmp_path / out)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    a, b = zr.read_index_dir(tmp_path / "ip"), zr.read_index_dir(tmp_path / "iz")
    assert [c["unique_kmers"] for c in a["manifest"]["colors"]] == \
        [c["unique_kmers"] for c in b["manifest"]["colors"]]
    for x, y in zip(a["banks"], b["banks"]):
        assert np.array_equal(x, y)
    bad = tmp_path / "bad.fa.xz"
    bad.write_bytes(lzma.compress(fa[0])[:-9])
    (tmp_path / "bad.fof").write_text(f"c\t{bad}\n")
    r = subprocess.run([str(cli), "build", "-f", str(tmp_path / "bad.fof"), "-o",
                        str(tmp_path / "ib")], capture_output=True, text=True)
    assert r.returncode != 0 and "xz" in r.stderr


def test_tuning_validation(tmp_path, small_colors):
    for bits, ok in ((4, True), (8, False)):
        idx = _build(small_colors[:2], fingerprint_bits=bits)
        out = tmp_path / f"idx{bits}"
        idx.write(str(out))
        with zindex_amd.Index.open(str(out)) as h:
            t = {"span": 1024, "tiles_per_wave": 1, "prefetch": 16, "waves_per_cu": 0,
                 "nontemporal": 0}
            if ok:
                h.tuning = t
                assert h.tuning["prefetch"] == 16
            else:
                with pytest.raises(zindex_amd.ZindexError):
                    h.tuning = t
            with pytest.raises(zindex_amd.ZindexError):
                h.tuning = dict(t, prefetch=3)


def test_modimizer_sampling_zero_is_rejected(small_colors):
    with pytest.raises(Exception, match="sampling"):
        _build(small_colors[:1], index_mode=2, modimizer_sampling=0)


@pytest.mark.parametrize("group,cap", [(1, 1 << 15), (2, 1 << 17), (1, 1 << 12)])
def test_hash_partitioned_build(tmp_path, small_colors, group, cap):
    """partition_large_subindexes (storage.rs:1525-1641): a size class over
    --max-subindex-size is rebuilt as hash partitions of 1/P of the slots; every
    key is stored in the partition partition_id picks (storage.rs:99-101)."""
    idx = _build(small_colors, fingerprint_bits=8, hashes=4, merge_group_size=group,
                 max_subindex_size=cap)
    man = idx.manifest()
    assert man["version"] == 9
    routed = [b for b in man["banks"] if b.get("routing")]
    source = routed[0]["routing"]["source_slots"]
    local, rb = source, (routed[0]["row_bits"] + 7) // 8
    while local * rb > cap:
        local //= 2
    parts = source // local
    assert parts > 1 and len(routed) == parts
    assert all(b["routing"]["partitions"] == parts for b in routed)
    for p, b in enumerate(routed):
        r = b["routing"]
        assert r["partition"] == p and b["slots"] * parts == r["source_slots"]
        assert b["file"] == f"bank-{r['source_slots'].bit_length() - 1}-part-{p}-of-{parts}.zorbank.zst"
        assert b["slots"] * ((b["row_bits"] + 7) // 8) <= cap
    # the partitions hold each group's keys exactly once between them
    for g in range(len(routed[0]["groups"])):
        assert len({tuple(b["groups"][g]["members"]) for b in routed}) == 1
    out = tmp_path / "idx"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    with zindex_amd.Index.open(str(out)) as reopened:
        assert reopened.meta["n_banks"] == parts and reopened.meta["n_classes"] == 1
    o = zr.OracleIndex(disk["manifest"], disk["banks"], disk["recovery"])
    total = 0
    for c, seq in enumerate(small_colors):
        keys = np.unique(zr.scan_kmers(seq, 31))
        unions = sum(b["groups"][[i for i, gg in enumerate(b["groups"]) if c in gg["members"]][0]]
                     ["union_cardinality"] for b in routed)
        if group == 1:
            assert unions == len(keys)
        scores = np.zeros(len(small_colors), np.uint64)
        for key in keys[:1500]:
            before = int(scores[c])
            o.lookup(int(key), scores)
            assert int(scores[c]) == before + 1, (c, int(key))
            total += 1
    assert total > 0


@pytest.mark.parametrize("bits,group,reverse", [(8, 1, False), (8, 3, True), (4, 2, False),
                                               (2, 5, True)])
def test_legacy_interleaved_layout_same_answers(tmp_path, small_colors, bits, group, reverse):
    """Legacy interleaved rows (fingerprint_first = false) answer like the plane
    layout (storage.rs:5343-5423): the same index rewritten interleaved, with
    its groups optionally reversed, gives the oracle the same counts, and the
    product's validation accepts it."""
    idx = _build(small_colors, fingerprint_bits=bits, hashes=4, merge_group_size=group,
                 slot_scale=0.8)
    plane, legacy = tmp_path / "plane", tmp_path / "legacy"
    idx.write(str(plane))
    zr.to_legacy_layout(plane, legacy,
                        (lambda gs: range(len(gs) - 1, -1, -1)) if reverse else None)
    with zindex_amd.Index.open(str(legacy)) as z:
        assert z.meta["n_colors"] == len(small_colors)
    dp, dl = zr.read_index_dir(plane), zr.read_index_dir(legacy)
    assert not dl["manifest"]["banks"][0]["fingerprint_first"]
    op = zr.OracleIndex(dp["manifest"], dp["banks"], dp["recovery"])
    ol = zr.OracleIndex(dl["manifest"], dl["banks"], dl["recovery"])
    rng = np.random.default_rng(3)
    reads = [c[i:i + 700] for c in small_colors for i in range(0, len(c) - 700, 3000)]
    reads += [_random_genome(rng, 700) for _ in range(10)]
    buf, off = zindex_amd.pack_reads(reads)
    tp, cp = op.query_reads(buf, off, threads=zr.cpu_count())
    tl, cl = ol.query_reads(buf, off, threads=zr.cpu_count())
    assert np.array_equal(tp, tl) and np.array_equal(cp, cl)
    assert cp.sum() > 0


def test_fusion_concatenates_indexes(tmp_path, small_colors):
    """fuse_many_indexes (fusion.rs:233-480): components 1..n, dataset ids of
    later inputs offset, recovery maps merged. The oracle's counts on the fused
    index are each input's counts side by side; the CLI and the product's
    validation accept the result."""
    parts = [small_colors[:2], small_colors[2:4], small_colors[4:]]
    dirs = []
    for i, cols in enumerate(parts):
        idx = _build(cols, fingerprint_bits=8, hashes=4, slot_scale=0.8,
                     merge_group_size=2 if i == 1 else 1)
        d = tmp_path / f"in{i}"
        idx.write(str(d))
        dirs.append(d)
    out = tmp_path / "fused"
    zindex_amd.fuse_many_indexes(dirs, out)
    with zindex_amd.Index.open(str(out)) as z:
        assert z.meta["version"] == 10 and z.meta["n_colors"] == len(small_colors)
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert [b["component"] for b in man["banks"]] == [1, 2, 3]
    assert all(b["file"].startswith(f"component-{b['component']}-") for b in man["banks"])
    assert man["source_document_ids"] == list(range(len(small_colors)))
    fused = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    rng = np.random.default_rng(4)
    reads = [c[i:i + 700] for c in small_colors for i in range(0, len(c) - 700, 4000)]
    reads += [_random_genome(rng, 700) for _ in range(8)]
    buf, off = zindex_amd.pack_reads(reads)
    tf, cf = fused.query_reads(buf, off, threads=zr.cpu_count())
    col = 0
    for d, cols in zip(dirs, parts):
        di = zr.read_index_dir(d)
        t, c = zr.OracleIndex(di["manifest"], di["banks"], di["recovery"]).query_reads(
            buf, off, threads=zr.cpu_count())
        assert np.array_equal(t, tf)
        assert np.array_equal(c, cf[:, col:col + len(cols)])
        col += len(cols)
    # CLI, and refusal to overwrite
    cli = str(ROOT / "zindex_amd" / "zorindex")
    r = subprocess.run([cli, "fusion", "-i", str(dirs[0]), str(dirs[1]), "-o",
                        str(tmp_path / "f2")], capture_output=True)
    assert r.returncode == 0, r.stderr
    with zindex_amd.Index.open(str(tmp_path / "f2")) as z:
        assert z.meta["n_colors"] == 4
    with pytest.raises(zindex_amd.ZindexError, match="already exists"):
        zindex_amd.fuse_indexes(dirs[0], dirs[1], out)


def test_fusion_rejects_incompatible_indexes(tmp_path, small_colors):
    a = _build(small_colors[:2], fingerprint_bits=8, hashes=4)
    b = _build(small_colors[2:4], fingerprint_bits=4, hashes=4)
    a.write(str(tmp_path / "a"))
    b.write(str(tmp_path / "b"))
    with pytest.raises(zindex_amd.ZindexError, match="fingerprint_bits"):
        zindex_amd.fuse_indexes(tmp_path / "a", tmp_path / "b", tmp_path / "out")


def _v1_case(tmp_path, small_colors, group=2):
    """A version-1 index (no fixed segment length, interleaved banks, fallback
    keys in the bank streams) whose fallback lists hold k-mers of the reads."""
    idx = _build(small_colors, fingerprint_bits=8, hashes=4, merge_group_size=group,
                 segment_length=0, slot_scale=0.9, drop_abandoned=1)
    plane = tmp_path / "plane"
    idx.write(str(plane))
    man = idx.manifest()
    rng = np.random.default_rng(5)
    reads = [c[i:i + 600] for c in small_colors for i in range(0, len(c) - 600, 5000)]
    keys = np.unique(np.concatenate([zr.scan_kmers(r, 31) for r in reads]))
    fallback = {}
    for key in rng.choice(keys, 300, replace=False):
        bi = int(rng.integers(len(man["banks"])))
        gi = int(rng.integers(len(man["banks"][bi]["groups"])))
        n = len(man["banks"][bi]["groups"][gi]["members"])
        mask = int(rng.integers(1, 1 << n))
        fallback.setdefault((bi, gi), []).append((int(key), mask))
    v1 = tmp_path / "v1"
    zr.to_v1(plane, v1, fallback)
    return plane, v1, reads, fallback


def test_version1_index_reads_fallback_keys(tmp_path, small_colors):
    """Version 1 (storage.rs:2555-2612, 2665-2674): the recovery map comes from
    the fallback lists inside the bank files. The oracle on the v1 files equals
    the oracle on the plane files plus a recovery map made of the same keys."""
    plane, v1, reads, fallback = _v1_case(tmp_path, small_colors)
    with zindex_amd.Index.open(str(v1)) as z:
        assert z.meta["version"] == 1 and z.meta["segment_length"] == 0
        assert len(z.bank_payload(0)) > 0
    dp, dl = zr.read_index_dir(plane), zr.read_index_dir(v1)
    man = dp["manifest"]
    by = {}
    for (bi, gi), entries in fallback.items():
        members = man["banks"][bi]["groups"][gi]["members"]
        for key, mask in entries:
            h = zr.oracle().zo_abandoned_hash(key, man["seed"])
            by.setdefault(h, set()).update(members[i] for i in range(len(members)) if mask >> i & 1)
    hashes = sorted(by)
    off = np.cumsum([0] + [len(by[h]) for h in hashes]).astype(np.uint64)
    origins = np.array([o for h in hashes for o in sorted(by[h])], np.uint32)
    rec = (np.array(hashes, np.uint32), off, origins)
    op = zr.OracleIndex(man, dp["banks"], rec)
    ol = zr.OracleIndex(dl["manifest"], dl["banks"], dl["recovery"])
    buf, off = zindex_amd.pack_reads(reads)
    tp, cp = op.query_reads(buf, off, threads=zr.cpu_count())
    tl, cl = ol.query_reads(buf, off, threads=zr.cpu_count())
    assert np.array_equal(tp, tl) and np.array_equal(cp, cl)
    op0 = zr.OracleIndex(man, dp["banks"], None)
    assert (cl >= op0.query_reads(buf, off, threads=zr.cpu_count())[1]).all()
    assert cl.sum() > op0.query_reads(buf, off, threads=zr.cpu_count())[1].sum()


def test_version1_bad_fallback_fingerprint_is_rejected(tmp_path, small_colors):
    plane, v1, reads, fallback = _v1_case(tmp_path, small_colors)
    man = json.loads((v1 / "manifest.json").read_text())
    bank = man["banks"][0]
    raw = bytearray(zr.zstd_decompress((v1 / bank["file"]).read_bytes()))
    n = 40 + bank["slots"] * ((bank["row_bits"] + 7) // 8)
    # the first group with a fallback key: flip its value's fingerprint
    at = n
    for g in bank["groups"]:
        cnt = int.from_bytes(raw[at:at + 8], "little")
        at += 8
        if cnt:
            raw[at + 8] ^= 1
            break
    (v1 / bank["file"]).write_bytes(zr.zstd_compress(bytes(raw)))
    with zindex_amd.Index.open(str(v1)) as z:
        with pytest.raises(zindex_amd.ZindexError, match="fallback fingerprint"):
            z.load()


def test_partition_cap_too_small_is_rejected(small_colors):
    with pytest.raises(zindex_amd.ZindexError, match="too small"):
        _build(small_colors[:2], fingerprint_bits=8, hashes=4, max_subindex_size=1 << 8)


def test_false_positive_rate_within_bound(small_colors):
    bits = 4
    idx = _build(small_colors, fingerprint_bits=bits, hashes=4)
    man = json.loads(json.dumps(idx_manifest(idx)))
    payloads = [idx.bank_payload(i) for i in range(len(man["banks"]))]
    o = zr.OracleIndex(man, payloads, None)
    present = set()
    for seq in small_colors:
        present.update(zr.scan_kmers(seq, 31).tolist())
    rng = np.random.default_rng(11)
    probes = [int(x) for x in rng.integers(0, 1 << 62, 20000, dtype=np.uint64)
              if int(x) not in present]
    scores = np.zeros(len(small_colors), np.uint64)
    for key in probes:
        o.lookup(key, scores)
    n = len(probes) * len(small_colors)
    p = 2.0 ** -bits
    hits = int(scores.sum())
    # option_coverage_cli.rs:247-290: within 8 sigma of the binomial mean
    assert abs(hits - n * p) <= 8 * (n * p * (1 - p)) ** 0.5


def idx_manifest(idx):
    """Manifest of an in-memory index, via a temporary write."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "i"
        idx.write(str(out))
        return json.loads((out / "manifest.json").read_text())


def test_high_load_abandons_into_recovery_and_stays_exact(tmp_path):
    rng = np.random.default_rng(5)
    colors = [_random_genome(rng, 60000) for _ in range(3)]
    # slot_scale 0.5 over-fills the filter: peeling must abandon keys
    idx = _build(colors, fingerprint_bits=8, hashes=4, slot_scale=0.5)
    out = tmp_path / "hl"
    idx.write(str(out))
    disk = zr.read_index_dir(out)
    man = disk["manifest"]
    assert man["abandoned_hashes"] > 0 and disk["recovery"] is not None
    assert sum(g["abandoned_keys"] for b in man["banks"] for g in b["groups"]) > 0
    o = zr.OracleIndex(man, disk["banks"], disk["recovery"])
    for c, seq in enumerate(colors):
        scores = np.zeros(len(colors), np.uint64)
        keys = np.unique(zr.scan_kmers(seq, 31))
        for key in keys:
            o.lookup(int(key), scores)
        assert int(scores[c]) == len(keys)
    # without recovery, stored keys go missing
    o2 = zr.OracleIndex(man, disk["banks"], disk["recovery"], recover=False)
    scores = np.zeros(len(colors), np.uint64)
    keys = np.unique(zr.scan_kmers(colors[0], 31))
    for key in keys:
        o2.lookup(int(key), scores)
    assert int(scores[0]) < len(keys)


@pytest.fixture(scope="module")
def readme_index(tmp_path_factory):
    d = tmp_path_factory.mktemp("readme")
    fof = d / "three.fof"
    names = list(vectors()["readme_bacteria"]["unique_kmers_k31"])
    fof.write_text("".join(f"{n}\t{GOLDEN / 'bacteria' / 'genomes' / (n + '.fna.gz')}\n"
                           for n in names))
    out = d / "idx"
    zindex_amd.build(str(fof), str(out), k=31)
    return out


def test_readme_unique_kmers_known_answers(readme_index):
    man = json.loads((readme_index / "manifest.json").read_text())
    expected = vectors()["readme_bacteria"]["unique_kmers_k31"]
    assert {c["name"]: c["unique_kmers"] for c in man["colors"]} == expected


def test_readme_query_known_answer_on_oracle(readme_index):
    v = vectors()["readme_bacteria"]
    o = zr.OracleIndex.from_dir(readme_index)
    recs = zr.read_fasta_records(GOLDEN / "bacteria" / "query.fa")
    buf, off = zindex_amd.pack_reads([s for _, s in recs])
    totals, counts = o.query_reads(buf, off)
    names = [c["name"] for c in o.manifest["colors"]]
    src = names.index(v["query_fa_source"])
    assert int(totals.sum()) == v["query_fa_total"]
    assert int(counts[:, src].sum()) == v["query_fa_matches"]


def test_info_command_runs_on_cpu(readme_index, tmp_path):
    out = tmp_path / "info.txt"
    fd = os.open(out, os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
    try:
        zindex_amd.info_command(str(readme_index), out_fd=fd)
    finally:
        os.close(fd)
    text = out.read_text()
    assert "indexed_documents\t3\n" in text and "fingerprint_bits\t8\n" in text
    assert text.count("bank-") >= 1


# the key lines of info_command in order (storage.rs:3914-4060), for a
# non-fused, non-findere index without anchors, segment length or limits
INFO_KEYS = [
    "collection_status", "indexed_documents", "index", "k", "hashes", "seed", "fingerprint_bits",
    "index_mode", "technical_layout", "colors", "groups", "banks", "size_classes",
    "max_subindex_size", "largest_subindex_bytes", "hash_partitions", "input_documents",
    "skipped_documents", "memory_guard_split_groups", "minimizer_size", "modimizer_sampling",
    "minimizer_scheme", "homopolymer_compression", "seed_pattern", "randstrobemers",
    "merge_planner", "planning_pair_evaluations", "planning_possible_pair_evaluations",
    "merge_reverts_enabled", "construction_cell_storage", "construction_abandoned_storage",
    "segment_length", "overflow_groups", "merge_column_order", "column_order_method",
    "abandoned_recovery",
    "abandoned_hash_bits", "abandoned_keys", "abandoned_unique_hashes",
    "abandoned_dataset_associations", "abandoned_map_ram_bytes", "logical_bits", "padding_bits",
    "stack_bytes", "stack_stored_bytes", "fallback_bytes", "structure_bytes",
    "singleton_logical_bits", "predicted_logical_bits", "planned_groups", "rejected_groups",
    "exact_group_verifications", "sketch_seconds", "plan_seconds", "build_seconds",
    "metadata_bytes", "uncompressed_index_bytes", "disk_bytes", "min_subindex_documents",
    "promoted_groups", "max_subindexes"]


def test_info_lines_follow_the_reference_order(readme_index, tmp_path):
    out = tmp_path / "info.txt"
    fd = os.open(out, os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
    try:
        zindex_amd.info_command(str(readme_index), out_fd=fd)
    finally:
        os.close(fd)
    lines = out.read_text().splitlines()
    head = [l.split("\t")[0] for l in lines[:lines.index(
        "bank\tslots\tgroups\trow_bits\tpadding_bits_per_row\tfallback_keys\tdocuments")]]
    expected = [k for k in INFO_KEYS if k in head]
    assert head == expected
    man = json.loads((readme_index / "manifest.json").read_text())
    optional = {"segment_length", "max_subindexes", "column_order_method"}
    assert set(INFO_KEYS) - set(head) <= optional
    disk = sum(p.stat().st_size for p in readme_index.iterdir())
    assert f"disk_bytes\t{disk}" in lines
    assert f"input_documents\t{len(man['colors'])}" in lines


def _tamper(src: Path, dst: Path, edit):
    shutil.copytree(src, dst)
    man = json.loads((dst / "manifest.json").read_text())
    edit(man)
    (dst / "manifest.json").write_text(json.dumps(man))


@pytest.mark.parametrize("name,edit", [
    ("version", lambda m: m.update(version=11)),
    ("hashes", lambda m: m.update(hashes=1)),
    ("bits", lambda m: m.update(fingerprint_bits=3)),
    ("segment", lambda m: m.update(segment_length=None)),
    ("width", lambda m: m["banks"][0]["groups"][0].update(width=9)),
    ("offset", lambda m: m["banks"][0]["groups"][1].update(offset=0)),
    ("member", lambda m: m["banks"][0]["groups"][0].update(members=[99])),
    ("dup", lambda m: m["banks"][0]["groups"][1].update(members=[0])),
    ("file", lambda m: m["banks"][0].update(file="bank-3.zorbank.zst")),
    ("slots", lambda m: m["banks"][0].update(slots=3)),
    ("recovery", lambda m: m.update(abandoned_file=None)),
    ("abandoned_keys", lambda m: m["banks"][0]["groups"][0].pop("abandoned_keys")),
    ("missing", lambda m: m.pop("seed")),
    ("backend", lambda m: m.update(backend="xor")),
    ("minimizer", lambda m: m.update(index_mode=1, k=48)),
])
def test_manifest_validation_rejects(readme_index, tmp_path, name, edit):
    bad = tmp_path / name
    _tamper(readme_index, bad, edit)
    with pytest.raises(zindex_amd.ZindexError) as e:
        zindex_amd.Index.open(str(bad))
    assert e.value.code == -3


def test_truncated_bank_is_rejected(readme_index, tmp_path):
    bad = tmp_path / "trunc"
    shutil.copytree(readme_index, bad)
    bank = next(bad.glob("bank-*.zorbank.zst"))
    data = bank.read_bytes()
    bank.write_bytes(data[: len(data) // 2])
    fd = os.open(os.devnull, os.O_WRONLY)
    try:
        with pytest.raises(zindex_amd.ZindexError):
            zindex_amd.info_command(str(bad), out_fd=fd)
    finally:
        os.close(fd)


def test_cli_build_and_info(tmp_path):
    cli = ROOT / "zindex_amd" / "zorindex"
    fof = tmp_path / "f.fof"
    rng = np.random.default_rng(2)
    paths = []
    for i in range(3):
        p = tmp_path / f"g{i}.fa"
        p.write_bytes(b">r\n" + _random_genome(rng, 5000) + b"\n")
        paths.append(p)
    fof.write_text("".join(f"{p}\n" for p in paths))
    out = tmp_path / "cli_idx"
    r = subprocess.run([str(cli), "build", "-f", str(fof), "-o", str(out), "-b", "16", "-H", "3"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    man = json.loads((out / "manifest.json").read_text())
    assert [c["name"] for c in man["colors"]] == ["g0", "g1", "g2"]
    r = subprocess.run([str(cli), "info", "-i", str(out)], capture_output=True, text=True)
    assert r.returncode == 0 and "hashes\t3" in r.stdout
    r = subprocess.run([str(cli), "build", "-f", str(fof), "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "exists" in r.stderr
    # seed options (src/bin/zorindex.rs:86-111)
    for name, args, fields in [
            ("fd", ["-i", "findere", "-z", "8", "-c"],
             {"index_mode": 3, "findere_z": 8, "homopolymer_compression": True}),
            ("sp", ["-i", "kmers", "-p", K31_PATTERN, "-d", "4"],
             {"index_mode": 0, "seed_pattern": K31_PATTERN, "modimizer_sampling": 4}),
            ("md", ["-i", "modimizers", "-d", "4"], {"index_mode": 2, "modimizer_sampling": 4}),
            ("rs", ["-i", "kmers", "-r", "-l", "7", "-u", "40"],
             {"randstrobemers": {"min_distance": 7, "max_distance": 40}})]:
        o = tmp_path / name
        r = subprocess.run([str(cli), "build", "-f", str(fof), "-o", str(o)] + args,
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        man = json.loads((o / "manifest.json").read_text())
        assert all(man[k] == v for k, v in fields.items()), (name, man)
    r = subprocess.run([str(cli), "build", "-f", str(fof), "-o", str(tmp_path / "bad"), "-p",
                        "##_"], capture_output=True, text=True)
    assert r.returncode != 0 and "symmetric" in r.stderr
