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
    """Binomial(len, rate) uniformly placed bases each become one of the three
    other bases (a repeated position is substituted once)."""
    out = seq.copy()
    n = int(rng.binomial(len(out), rate)) if len(out) else 0
    if n:
        pos = rng.integers(0, len(out), n)
        shift = rng.integers(1, 4, n, dtype=np.uint8)
        out[pos] = ACGT[(_CODE[seq[pos]] + shift) % 4]
    return out


def indels(rng: np.random.Generator, seq: np.ndarray, rate: float, max_len: int = 10) -> np.ndarray:
    """Insert or delete runs of 1..max_len bases at a fraction `rate` of
    positions (half insertions of random bases, half deletions)."""
    n_events = int(rng.binomial(len(seq), rate)) if len(seq) else 0
    if n_events == 0:
        return seq.copy()
    pos = np.unique(rng.integers(0, len(seq), n_events))
    lens = rng.integers(1, max_len + 1, len(pos))
    ins = rng.random(len(pos)) < 0.5
    dp, dl = pos[~ins], lens[~ins]
    deleted = np.zeros(0, dtype=np.int64)
    if len(dp):
        starts = np.repeat(np.cumsum(dl) - dl, dl)
        idx = np.repeat(dp, dl) + (np.arange(int(dl.sum())) - starts)
        deleted = np.unique(idx[idx < len(seq)])
    kept = np.delete(seq, deleted)
    ip, il = pos[ins], lens[ins]
    if len(ip) == 0:
        return kept
    # insertion before original index ip lands after the kept bases below ip
    before = ip - np.searchsorted(deleted, ip)
    return np.insert(kept, np.repeat(before, il), random_genome(rng, int(il.sum())))


def substitution_tree(rng: np.random.Generator, root: np.ndarray, leaves: int,
                      sub_range=(0.005, 0.05), indel_rate=0.01):
    """Yield (leaf id, sequence) for the `leaves` colours (a power of two) of a
    balanced binary tree. Each leaf's total substitution rate is drawn
    uniformly from sub_range and spread evenly over the edges of its path
    (an edge carries the mean rate of the leaves below it), so siblings share
    most substitutions. Indels (lengths 1..10) hit `indel_rate` of positions
    of each leaf, drawn at the leaf. Generated depth first: memory O(depth)."""
    depth = int(round(np.log2(leaves)))
    assert 1 << depth == leaves
    totals = rng.uniform(sub_range[0], sub_range[1], leaves)
    per_edge = totals / max(depth, 1)
    seeds = rng.integers(0, 2**62, size=2 * leaves, dtype=np.int64)

    def walk(node: int, level: int, seq: np.ndarray):
        if level == depth:
            leaf = node - leaves
            s = seq
            if indel_rate:
                s = indels(np.random.default_rng(int(seeds[node]) ^ 0x5EED), s, indel_rate)
            yield leaf, s
            return
        for child in (2 * node, 2 * node + 1):
            crng = np.random.default_rng(int(seeds[child]))
            span = 1 << (depth - level - 1)
            lo = child * span - leaves
            rate = float(per_edge[lo: lo + span].mean())
            yield from walk(child, level + 1, substitute(crng, seq, rate))

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


def sample_reads_fast(rng: np.random.Generator, colors: list, n: int, lengths: np.ndarray,
                      error: float, rc_fraction: float = 0.5):
    """As sample_reads, with the substitution errors drawn once over the whole
    buffer (the per-read loop only copies slices)."""
    src = rng.integers(0, len(colors), n)
    clen = np.array([len(c) for c in colors], dtype=np.int64)
    ln = np.minimum(np.asarray(lengths, dtype=np.int64), clen[src])
    starts = (rng.random(n) * (clen[src] - ln + 1)).astype(np.int64)
    rc = rng.random(n) < rc_fraction
    off = np.zeros(n + 1, dtype=np.uint64)
    off[1:] = np.cumsum(ln)
    buf = np.empty(int(off[-1]), dtype=np.uint8)
    o = off.astype(np.int64)
    for i in range(n):
        piece = colors[src[i]][starts[i]:starts[i] + ln[i]]
        buf[o[i]:o[i + 1]] = _COMP[piece[::-1]] if rc[i] else piece
    if error and len(buf):
        buf = substitute(rng, buf, error)
    return buf, off, src


def lognormal_lengths(rng: np.random.Generator, n: int, median: float, sigma: float,
                      lo: int, hi: int) -> np.ndarray:
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, n)), lo, hi).astype(np.int64)
