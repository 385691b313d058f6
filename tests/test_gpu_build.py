"""The GPU peeling build (zg_build_config.device = 1): valid ZOR filters in the
reference format. Checked against the CPU build of the same datasets: the same
abandoned keys (the 2-core left by peeling is unique and broken by the same
rule on the host), no false negatives on stored k-mers with recovery, query
parity with the oracle on the GPU-built index, and false-positive rates
within the 2^-b bound."""
import numpy as np
import pytest

import zindex_ref as zr
from test_gpu_parity import build, check_parity, make_colors, make_reads, oracle_for, say

zindex_amd = pytest.importorskip("zindex_amd")

pytestmark = pytest.mark.gpu


def no_false_negatives(idx, colors):
    """Every colour's own sequence, queried as reads of 2 kbp, matches that
    colour on every k-mer (recovery on)."""
    for c, seq in enumerate(colors):
        reads = [seq[i:i + 2000] for i in range(0, len(seq), 2000)]
        totals, counts, _ = idx.query_reads(reads, records=True)
        assert np.array_equal(counts[:, c].astype(np.uint64), totals), f"colour {c}"


@pytest.mark.parametrize("bits,hashes,scale,group", [
    (8, 4, 1.0, 1), (4, 3, 1.0, 1), (16, 8, 1.0, 1), (1, 4, 1.0, 1), (32, 4, 1.0, 1),
    (8, 2, 1.0, 1), (8, 4, 0.55, 1), (8, 3, 0.6, 3), (2, 4, 1.0, 4)])
def test_gpu_peel_build(bits, hashes, scale, group):
    colors = make_colors(90 + bits + hashes, 6, 30000)
    cpu = build(colors, fingerprint_bits=bits, hashes=hashes, slot_scale=scale,
                merge_group_size=group)
    gpu = build(colors, fingerprint_bits=bits, hashes=hashes, slot_scale=scale,
                merge_group_size=group, device=1)
    mc, mg = cpu.meta, gpu.meta
    say(f"b{bits} H{hashes} scale {scale} group {group}: abandoned {mc['abandoned_origins']} "
        f"(cpu) {mg['abandoned_origins']} (gpu)")
    assert mc["abandoned_origins"] == mg["abandoned_origins"]
    hc, oc, gc = cpu.recovery()
    hg, og, gg = gpu.recovery()
    assert np.array_equal(hc, hg) and np.array_equal(oc, og) and np.array_equal(gc, gg)
    assert cpu.manifest()["banks"] == gpu.manifest()["banks"]
    gpu.load()
    no_false_negatives(gpu, colors)
    reads = make_reads(91 + bits, colors, 150)
    check_parity(gpu, oracle_for(gpu), reads, label=f"gpu-built b{bits} H{hashes}")


@pytest.mark.parametrize("bits", [1, 4, 8])
def test_gpu_peel_false_positive_rate(bits):
    """Random sequence shares no k-mer with the colours: the hit rate per colour
    is the filter's false-positive rate, 2^-b within the tolerance
    tests/option_coverage_cli.rs:247-290 of the reference allows."""
    colors = make_colors(95 + bits, 4, 60000)
    idx = build(colors, fingerprint_bits=bits, hashes=4, device=1)
    idx.load()
    # reads from a stream no colour was drawn from, checked to share no k-mer
    rng = np.random.default_rng(10_000 + bits)
    reads = [zindex_amd.synth.random_genome(rng, 2000).tobytes() for _ in range(100)]
    present = set()
    for c in colors:
        present.update(zr.scan_kmers(c, 31).tolist())
    assert not any(set(zr.scan_kmers(r, 31).tolist()) & present for r in reads)
    totals, counts, agg = idx.query_reads(reads, records=False, recover=False)
    n = int(totals.sum())
    p = 2.0 ** -bits
    # the reference's tolerance: |hits - n p| <= 8 sqrt(n p (1 - p)) + 8
    bound = p + (8 * np.sqrt(n * p * (1 - p)) + 8) / n
    rates = agg / n
    say(f"b{bits}: {n} k-mers, FPR {rates.tolist()} bound {bound:.5f}")
    assert np.all(rates <= bound)
