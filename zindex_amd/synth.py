"""Seeded synthetic DNA for benchmarks and tests: random genomes, substitution
and indel descendants, a balanced substitution tree of colours, and reads
sampled from colours with errors and random strand.

Sequences are uint8 arrays of ASCII ACGT.
"""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
_CODE = np.full(256, 255, dtype=np.uint8)
_CODE[ACGT] = np.arange(4, dtype=np.uint8)
_COMP = np.arange(256, dtype=np.uint8)
_COMP[ACGT] = ACGT[::-1]


def random_genome(rng: np.random.Generator, n: int) -> np.ndarray:
    return ACGT[rng.integers(0, 4, n, dtype=np.uint8)]


def substitute(rng: np.random.Generator, seq: np.ndarray, rate: float) -> np.ndarray:
    """Each base becomes one of the three other bases with probability rate."""
    out = seq.copy()
    pos = np.nonzero(rng.random(len(out)) < rate)[0]
    shift = rng.integers(1, 4, len(pos), dtype=np.uint8)
    out[pos] = ACGT[(_CODE[out[pos]] + shift) % 4]
    return out


def indels(rng: np.random.Generator, seq: np.ndarray, rate: float, max_len: int = 10) -> np.ndarray:
    """Insert or delete runs of 1..max_len bases at a fraction `rate` of positions."""
    n_events = rng.binomial(len(seq), rate)
    if n_events == 0:
        return seq.copy()
    pos = np.sort(rng.choice(len(seq), n_events, replace=False))
    lens = rng.integers(1, max_len + 1, n_events)
    kind = rng.random(n_events) < 0.5  # True = insertion
    pieces, at = [], 0
    for p, ln, ins in zip(pos.tolist(), lens.tolist(), kind.tolist()):
        if p < at:
            continue
        pieces.append(seq[at:p])
        if ins:
            pieces.append(random_genome(rng, ln))
            at = p
        else:
            at = p + ln
    pieces.append(seq[at:])
    return np.concatenate(pieces)


def substitution_tree(rng: np.random.Generator, root: np.ndarray, leaves: int,
                      sub_range=(0.005, 0.05), indel_rate=0.01):
    """Yield `leaves` colour sequences (a power of two) of a balanced binary
    tree: each leaf's total substitution rate is drawn from sub_range and
    spread evenly over the depth of its path; indels hit `indel_rate` of
    positions over the path. Internal nodes are shared, so siblings are
    closer than cousins. Generated depth first: memory is O(depth)."""
    depth = int(round(np.log2(leaves)))
    assert 1 << depth == leaves
    totals = rng.uniform(sub_range[0], sub_range[1], leaves)
    per_edge = totals / max(depth, 1)
    per_edge_indel = indel_rate / max(depth, 1)
    seeds = rng.integers(0, 2**63, size=2 * leaves, dtype=np.int64)

    def walk(node: int, level: int, seq: np.ndarray):
        if level == depth:
            yield node - leaves, seq
            return
        for child in (2 * node, 2 * node + 1):
            crng = np.random.default_rng(int(seeds[child % len(seeds)]) ^ child)
            # edge rate: mean of the leaves below this child
            lo = child << (depth - level - 1)
            rate = float(per_edge[lo - leaves: lo - leaves + (1 << (depth - level - 1))].mean())
            s = substitute(crng, seq, rate)
            if per_edge_indel:
                s = indels(crng, s, per_edge_indel)
            yield from walk(child, level + 1, s)

    yield from walk(1, 0, root)


def reverse_complement(seq: np.ndarray) -> np.ndarray:
    return _COMP[seq[::-1]]


def sample_reads(rng: np.random.Generator, colors: list, n: int, lengths: np.ndarray,
                 error: float, rc_fraction: float = 0.5):
    """Reads from random colours/offsets with substitution errors and random
    strand. Returns (buffer, offsets[n+1], source colour ids)."""
    src = rng.integers(0, len(colors), n)
    parts, off = [], np.zeros(n + 1, dtype=np.uint64)
    for i in range(n):
        c = colors[src[i]]
        ln = int(min(lengths[i], len(c)))
        start = int(rng.integers(0, len(c) - ln + 1))
        read = c[start:start + ln]
        if error:
            read = substitute(rng, read, error)
        if rng.random() < rc_fraction:
            read = reverse_complement(read)
        parts.append(read)
        off[i + 1] = off[i] + ln
    buf = np.concatenate(parts) if parts else np.zeros(0, np.uint8)
    return buf, off, src


def lognormal_lengths(rng: np.random.Generator, n: int, median: float, sigma: float,
                      lo: int, hi: int) -> np.ndarray:
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, n)), lo, hi).astype(np.int64)
