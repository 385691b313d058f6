"""The allocation-free query entry point (zg_query_reads_noalloc, a static
device array per GPU instead of hipMalloc) against zg_query_reads through the
C ABI: bit-identical per-read totals, per-read per-colour counts and
aggregates on two calls in a row, the second with fewer reads and bases (so
stale contents of the shared array would show), for dense k-mers, merged
groups, minimizers, modimizers and a colour count whose partial sums narrow
the grid. The first call is also checked against the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

import zindex_ref as zr

zindex_amd = pytest.importorskip("zindex_amd")
from zindex_amd import _lib  # noqa: E402
from test_gpu_parity import build, make_colors, make_reads, oracle_for, say  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = {
    "kmers b8 h4": (dict(fingerprint_bits=8, hashes=4), 12, 30000),
    "kmers b4 h3 groups of 4": (dict(fingerprint_bits=4, hashes=3, merge_group_size=4), 20, 20000),
    "minimizers k31 m19": (dict(fingerprint_bits=8, hashes=3, index_mode=1, k=31,
                                minimizer_size=19), 10, 30000),
    "modimizers 1/3": (dict(fingerprint_bits=8, hashes=3, index_mode=2, modimizer_sampling=3),
                       10, 30000),
    "1100 colours": (dict(fingerprint_bits=8, hashes=4), 1100, 3000),
}


def both(idx, reads, records, recover):
    buf, off = zindex_amd.pack_reads(reads)
    return (idx.query_reads(buf, off, records=records, recover=recover),
            idx.query_reads_noalloc(buf, off, records=records, recover=recover))


def assert_same(want, got, label):
    (t0, c0, a0), (t1, c1, a1) = want, got
    assert np.array_equal(t0, t1), label
    if c0 is not None:
        bad = np.argwhere(c0 != c1)
        assert len(bad) == 0, f"{label}: {len(bad)} count mismatches, first {bad[:5].tolist()}"
    assert np.array_equal(a0, a1), label


@pytest.mark.parametrize("case", list(CASES))
def test_noalloc_matches_query_reads(case):
    cfg, n_colors, length = CASES[case]
    colors = make_colors(700 + n_colors, n_colors, length)
    idx = build(colors, **cfg)
    idx.load()
    big = make_reads(701, colors, 400, lo=20, hi=2500)
    small = make_reads(702, colors, 37, lo=20, hi=300)
    for records in (True, False):
        for recover in (True, False):
            label = f"{case} records={records} recover={recover}"
            # the two allocation-free calls back to back, then the references
            buf_b, off_b = zindex_amd.pack_reads(big)
            buf_s, off_s = zindex_amd.pack_reads(small)
            say(f"noalloc {label}: {len(big)} reads / {int(off_b[-1])} bases, then "
                f"{len(small)} reads / {int(off_s[-1])} bases")
            got_b = idx.query_reads_noalloc(buf_b, off_b, records=records, recover=recover)
            got_s = idx.query_reads_noalloc(buf_s, off_s, records=records, recover=recover)
            want_b = idx.query_reads(buf_b, off_b, records=records, recover=recover)
            want_s = idx.query_reads(buf_s, off_s, records=records, recover=recover)
            assert_same(want_b, got_b, label + " (first call)")
            assert_same(want_s, got_s, label + " (second, smaller call)")
    # and the oracle on the first input
    orc = oracle_for(idx)
    buf, off = zindex_amd.pack_reads(big)
    exp_totals, exp_counts = orc.query_reads(buf, off, threads=zr.cpu_count())
    t, c, a = idx.query_reads_noalloc(buf, off, records=True)
    assert np.array_equal(t, exp_totals) and np.array_equal(c.astype(np.uint64), exp_counts)
    assert np.array_equal(a, exp_counts.sum(axis=0))
    assert exp_counts.sum() > 0


def test_noalloc_rejects_oversized_and_unsupported():
    colors = make_colors(710, 4, 20000)
    idx = build(colors, fingerprint_bits=8, hashes=3)
    idx.load()
    # 60 Mbp needs ~84 MiB of working buffers: refused before any GPU work
    reads = [b"ACGT" * 15_000_000]
    with pytest.raises(_lib.ZindexError) as e:
        idx.query_reads_noalloc(reads)
    assert e.value.code == _lib.ZG_ERR_INVALID and "64 MiB" in str(e.value)
    # the same call still works on the allocating path, and the entry point
    # keeps working after the refusal
    idx.query_reads(reads)
    small = make_reads(711, colors, 20)
    assert_same(*both(idx, small, True, True), "after refusal")
    # homopolymer-compressed index: not covered
    hpc = build(colors, fingerprint_bits=8, hashes=3, homopolymer_compression=1)
    hpc.load()
    with pytest.raises(_lib.ZindexError) as e:
        hpc.query_reads_noalloc(small)
    assert e.value.code == _lib.ZG_ERR_UNSUPPORTED
    # null arguments
    rc = _lib.lib().zg_query_reads_noalloc(idx._h, None, None, C.c_uint64(3), 0, None, None, None)
    assert rc == _lib.ZG_ERR_INVALID
