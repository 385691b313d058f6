"""Pins the CPU oracle (oracle/oracle.c) against the reference's own known
answers, its brute-force test oracles and mathematical invariants."""
import ctypes as C
import itertools

import numpy as np
import pytest

import zindex_ref as zr
from zindex_ref import oracle, parse_int, u32p, u64p, vectors

MASK64 = (1 << 64) - 1


def sm64(x):
    return oracle().zo_splitmix64(x & MASK64)


def test_splitmix64_golden():
    for x, expected in vectors()["splitmix64"]["cases"]:
        assert sm64(x) == parse_int(expected)


def test_mixsplit_relation():
    for key, seed in vectors()["mixsplit_relation"]["cases"]:
        key, seed = parse_int(key), parse_int(seed)
        assert oracle().zo_mixsplit(key, seed) == sm64((key + seed) & MASK64)


def test_encode_base_and_full_kmer():
    v = vectors()
    for base, expected in v["encode_base"]["cases"]:
        got = oracle().zo_encode_base(ord(base))
        assert got == (-1 if expected is None else expected)
    for seq, expected in v["encode_full_kmer"]["cases"]:
        out = C.c_uint64()
        ok = oracle().zo_encode_full_kmer(seq.encode(), len(seq), C.byref(out))
        assert (out.value if ok else None) == expected


def test_canonical_kmer_counts():
    for seq, k, expected in vectors()["canonical_kmer_counts"]["cases"]:
        assert len(zr.scan_kmers(seq.encode(), k)) == expected


def _rc(seq: bytes) -> bytes:
    # src/lib.rs:869-880: complement (upper case) ACGT, keep other bytes
    comp = {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A")}
    return bytes(comp.get(ord(chr(c).upper()), c) for c in reversed(seq))


def _sequence(n, salt=0):
    return bytes(b"ACGT"[sm64(i ^ salt) & 3] for i in range(n))


def test_canonical_encoded_brute_force():
    rng = np.random.default_rng(1)
    for k in (1, 2, 5, 16, 31, 32):
        for _ in range(50):
            v = int(rng.integers(0, 1 << (2 * k), dtype=np.uint64)) if k < 32 else \
                int(rng.integers(0, 2**63, dtype=np.uint64)) * 2 + 1
            bases = [(v >> (2 * (k - 1 - i))) & 3 for i in range(k)]
            rc = 0
            for b in reversed(bases):
                rc = (rc << 2) | (3 - b)
            assert oracle().zo_canonical_encoded(v, k) == min(v, rc)


def test_kmers_match_independent_window_folding():
    # src/lib.rs:821-868: N and lowercase handling, both orientations
    seq = bytearray(_sequence(2400))
    seq[80] = ord("N")
    seq[900] = ord("n")
    seq[300:360] = bytes(seq[300:360]).lower()
    seq = bytes(seq)
    base = 0x9e3779b185ebca87
    for k in (1, 31, 32, 33, 49):
        expected = []
        for i in range(len(seq) - k + 1):
            w = seq[i:i + k]
            codes = [b"ACGT".find(bytes([c]).upper()) for c in w]
            if min(codes) < 0:
                continue
            if k <= 32:
                f = 0
                for c in codes:
                    f = (f << 2) | c
                r = 0
                for c in reversed(codes):
                    r = (r << 2) | (c ^ 3)
                expected.append(min(f, r))
            else:
                f = 0
                for c in codes:
                    f = (f * base + c + 1) & MASK64
                r = 0
                for c in reversed(codes):
                    r = (r * base + 4 - c) & MASK64
                expected.append(sm64(min(f, r) ^ sm64(k ^ 0x6b6d65722d763031)))
        got = zr.scan_kmers(seq, k)
        assert got.tolist() == expected, k
        rc = zr.scan_kmers(_rc(seq), k)
        assert got.tolist() == rc[::-1].tolist(), k


OCM_SEQS = [
    _sequence(190),
    b"ACGT" * 48,
    b"A" * 190,
    b"AAAACCCCGGGGTTTT" * 12,
]


@pytest.mark.parametrize("seq_id", range(len(OCM_SEQS)))
def test_ocm_rolling_matches_bruteforce_and_rc(seq_id):
    # src/open_closed_minimizers.rs:88-128
    seq = OCM_SEQS[seq_id]
    rc = _rc(seq)
    for k in range(1, 66, 2):
        for m in range(1, min(k, 31) + 1):
            for seed in (0, 69):
                pos, key = zr.scan_ocm(seq, k, m, seed)
                bpos, bkey = zr.scan_ocm_bruteforce(seq, k, m, seed)
                assert pos.tolist() == bpos.tolist() and key.tolist() == bkey.tolist(), (k, m, seed)
                rpos, rkey = zr.scan_ocm(rc, k, m, seed)
                fwd = set(zip(pos.tolist(), key.tolist()))
                rev = {(len(seq) - m - int(p), int(x)) for p, x in zip(rpos, rkey)}
                assert fwd == rev, (k, m, seed)
                if (k + m) % 7 == 0:  # window-local selection on a sample of windows
                    indexed = set(key.tolist())
                    for s in range(0, len(seq) - k + 1, 13):
                        wp, wk = zr.scan_ocm(seq[s:s + k], k, m, seed)
                        assert len(wk) == 1 and int(wk[0]) in indexed


def test_ocm_splits_runs_and_rejects_even_windows():
    seq = _sequence(300, 5)
    broken = seq[:120] + b"N" + seq[121:]
    pos, key = zr.scan_ocm(broken, 49, 31, 69)
    a_pos, a_key = zr.scan_ocm(seq[:120], 49, 31, 69)
    b_pos, b_key = zr.scan_ocm(seq[121:], 49, 31, 69)
    assert key.tolist() == a_key.tolist() + b_key.tolist()
    assert pos.tolist() == a_pos.tolist() + [p + 121 for p in b_pos.tolist()]
    assert len(zr.scan_ocm(seq, 48, 31, 69)[0]) == 0
    assert len(zr.scan_ocm(seq, 49, 32, 69)[0]) == 0


def test_homopolymer_compress_known_answer():
    # tests/minimizer_schemes_cli.rs:250-254
    assert zr.homopolymer_compress(b"AaAACccNNNTttAA") == b"ACNNNTA"
    assert zr.homopolymer_compress(b"") == b""
    assert zr.homopolymer_compress(b"acgtNNacgt") == b"ACGTNNACGT"


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_homopolymer_runs_give_the_compressed_features(mode):
    # tests/minimizer_schemes_cli.rs:255-282: two copies of ACGTA behind runs of
    # four or five equal bases, split by an N run
    expanded = b"AAAACCCCGGGGTTTTAAAAANNNAAAACCCCGGGGTTTTAAAAA"
    got = zr.read_features(expanded, 5, 69, mode=mode, m=3, sampling=1, hpc=True).tolist()
    once = zr.read_features(b"ACGTA", 5, 69, mode=mode, m=3, sampling=1).tolist()
    assert got == once * 2 and len(once) > 0


def _pext(x, mask):
    out, o = 0, 0
    for b in range(64):
        if (mask >> b) & 1:
            out |= ((x >> b) & 1) << o
            o += 1
    return out


def test_spaced_seed_projection_known_answers():
    # src/anchors.rs:750-773
    width, mask = zr.spaced_seed("##_##_##")
    assert width == 8 and bin(mask).count("1") // 2 == 6
    enc = _encode
    assert oracle().zo_spaced_project(enc(b"ACGTAACG"), width, mask) == enc(b"ACTACG")
    a, b = b"ACATACCG", b"ACGTAGCG"
    ca = oracle().zo_spaced_canonical(enc(a), width, mask)
    assert ca == oracle().zo_spaced_canonical(enc(b), width, mask)
    assert ca == oracle().zo_spaced_canonical(enc(_rc(a)), width, mask)
    # src/anchors.rs:776-791: projection is the bit extraction for every width
    for w in range(1, 32):
        pat = ["_"] * w
        for i in range(w):
            if i == 0 or sm64(min(i, w - 1 - i)) & 1:
                pat[i] = pat[w - 1 - i] = "#"
        width, mask = zr.spaced_seed("".join(pat))
        for i in range(100):
            key = sm64(i)
            assert oracle().zo_spaced_project(key, width, mask) == _pext(key, mask)


def _encode(s: bytes) -> int:
    v = 0
    for ch in s:
        v = (v << 2) | b"ACGT".index(ch)
    return v


@pytest.mark.parametrize("mode,k,m", [(0, 9, 3), (1, 31, 9)])
def test_spaced_seeds_mask_plain_features(mode, k, m):
    seq = _sequence(4000, 17)
    plain = zr.read_features(seq, k, 69, mode=mode, m=m)
    # src/anchors.rs:839-852: an all-retained pattern leaves keys unchanged
    full = "#" * (k if mode == 0 else m)
    assert zr.read_features(seq, k, 69, mode=mode, m=m, pattern=full).tolist() == plain.tolist()
    # src/anchors.rs:854-866: minimizers are selected before masking
    pat = "##__#__##" if mode == 1 else "##_#_#_##"
    width, mask = zr.spaced_seed(pat)
    got = zr.read_features(seq, k, 69, mode=mode, m=m, pattern=pat).tolist()
    assert got == [oracle().zo_spaced_canonical(int(x), width, mask) for x in plain]
    # src/anchors.rs:869-908: reverse complements keep the anchor multiset
    rc = zr.read_features(_rc(seq), k, 69, mode=mode, m=m, pattern=pat).tolist()
    assert sorted(rc) == sorted(got)


def test_strobe_hash_known_answer_and_pair_symmetry():
    # src/anchors.rs:823-838
    assert oracle().zo_strobe_hash(0x3332316f6c6c6568) == 0x2d119fced5ee39ce
    h = lambda s: oracle().zo_strobe_hash(min(_encode(s), _encode(_rc(s))))
    a, b = h(b"AC"), h(b"AG")
    assert a == h(b"GT") and b == h(b"CT")  # canonical strobes
    assert zr.link_strobes([0, 5], [a, b], 1, 10)[0] == a // 2 + b // 2
    assert zr.link_strobes([0, 5], [b, a], 1, 10)[0] == a // 2 + b // 2


@pytest.mark.parametrize("k,low,high,step,count,expected", [
    (11, 4, 12, 1, 61, 11859818861789259928), (11, 2, 5, 3, 21, 7111390358099897965),
    (21, 7, 15, 1, 51, 7250723734144174666)])
def test_linker_matches_upstream_strobealign_witnesses(k, low, high, step, count, expected):
    # src/anchors.rs:950-988: checksums of every key the upstream iterator emits
    seq = b"ACGTTGCACTGATCGATGGCATACCTAGGCTAACGTCGATCTTACGGAATCGCTAGCTGACCATGGTTACG"
    pos = list(range(0, len(seq) - k + 1, step))
    hashes = [oracle().zo_strobe_hash(int(zr.scan_kmers(seq[p:p + k], k)[0])) for p in pos]
    keys = zr.link_strobes(pos, hashes, low * step, high * step)
    assert len(keys) == count
    checksum = 0
    for key in keys:
        checksum = (checksum * 0x9e3779b185ebca87 + key) & MASK64
    assert checksum == expected


def test_randstrobe_distances_are_inclusive_with_self_fallback():
    # src/anchors.rs:1037-1048
    f = lambda s: zr.read_features(s, 3, 69, strobes=(5, 5))
    assert len(f(b"AC")) == 0 and len(f(b"ACGTACG")) == 5 and len(f(b"ACGTACGT")) == 6


@pytest.mark.parametrize("mode", [0, 1])
def test_randstrobes_respect_breaks_compression_and_orientation(mode):
    k = 9 if mode == 0 else 31
    kw = dict(mode=mode, m=9, pattern="##__#__##", strobes=(9, 40))
    seq = _sequence(300, 31)
    # src/anchors.rs:1013-1035
    once = zr.read_features(seq, k, 69, **kw).tolist()
    assert zr.read_features(seq + b"NNN" + seq, k, 69, **kw).tolist() == once * 2
    expanded = bytes(c for c in seq + b"NNN" + seq for _ in range(3))
    assert zr.read_features(expanded, k, 69, hpc=True, **kw).tolist() == \
        zr.read_features(zr.homopolymer_compress(expanded), k, 69, **kw).tolist()
    # src/anchors.rs:868-908, 990-1011: one anchor per plain feature in either
    # orientation; partner selection is directional
    long = _sequence(1000, 37)
    kw = dict(mode=mode, m=11, strobes=(15, 50))
    k = 11 if mode == 0 else 29
    fwd = zr.read_features(long, k, 69, **kw)
    rev = zr.read_features(_rc(long), k, 69, **kw)
    assert len(fwd) == len(rev) == len(zr.read_features(long, k, 69, mode=mode, m=11))
    assert sorted(fwd.tolist()) != sorted(rev.tolist())


def test_findere_consecutive_hits_in_one_colour():
    # src/findere.rs:466-477: six s-mers; colour 0 has runs of 2 and 3
    pos = [0]

    def hits(_key):
        h = [0, 0]
        h[1 if pos[0] == 2 else 0] = 1
        pos[0] += 1
        return h
    assert zr.findere_sequence(b"ACGTACGT", 5, 2, 2, hits) == ([1, 0], 4, 6)


def test_findere_short_records_and_breaks():
    # src/findere.rs:479-490
    positive = lambda _key: [1]
    for seq in (b"ACGT", b"ACGTNACGT"):
        assert zr.findere_sequence(seq, 5, 2, 1, positive) == ([0], 0, 0)
    assert zr.findere_sequence(b"ACGTANACGTA", 5, 2, 1, positive) == ([2], 2, 6)


def test_findere_equals_window_intersection():
    # src/findere.rs:493-529, reverse complement and a substitution included
    reference = b"ACGTTAGCGGATCACGTAC"
    queries = [reference, b"GTACGTGATCCGCTAACGT", b"AAACGTTAGCGGTTCACGTAC"]
    for z in (0, 1, 3, 4):
        k, s = 7, 7 - z
        smers = set(zr.scan_kmers(reference, s).tolist())
        for q in queries:
            got, total, _ = zr.findere_sequence(q, k, z, 1, lambda key: [int(key in smers)])
            expected = sum(all(x in smers for x in zr.scan_kmers(q[i:i + k], s).tolist())
                           for i in range(len(q) - k + 1))
            assert got == [expected] and total == len(q) - k + 1


@pytest.mark.parametrize("k,z", [(31, 10), (41, 10), (5, 2), (31, 0)])
def test_findere_brute_force_k_windows(k, z):
    # the synthetic hit function of src/findere.rs:280-285 against the k-window
    # rule: a window of k valid bases scores a colour when all its s-mers hit
    synthetic = lambda key: [1, int(key % 127 != 0), int(key % 7 != 0)]
    seq = bytearray(_sequence(1500, 23))
    for at in (100, 101, 700, 1300):
        seq[at] = ord("N")
    seq = bytes(seq[:400]) + b"ACGT" + bytes(seq[400:])
    s = k - z
    exp, total = [0, 0, 0], 0
    for i in range(len(seq) - k + 1):
        w = seq[i:i + k]
        if b"N" in w:
            continue
        total += 1
        keys = zr.scan_kmers(w, s).tolist()
        for c in range(3):
            exp[c] += all(synthetic(x)[c] for x in keys)
    got, got_total, _ = zr.findere_sequence(seq, k, z, 3, synthetic)
    assert got == exp and got_total == total


# src/lib.rs:1089-1113: with sampling 1 the modimizers are every canonical k-mer
@pytest.mark.parametrize("seq,k,seed", [(b"ACGTACGT", 3, 1), (b"AAAACCCC", 4, 2),
                                        (b"TTTTGGGG", 5, 3), (b"GATTACA", 3, 4),
                                        (b"ACGTNACGT", 4, 5), (b"CCCCCCCC", 2, 6)])
def test_modimizers_sampling_one_are_all_kmers(seq, k, seed):
    assert zr.read_features(seq, k, seed, mode=2, sampling=1).tolist() == \
        zr.scan_kmers(seq, k).tolist()


def test_modimizers_nest_and_thin_by_sampling():
    # sampling s keeps the k-mers whose splitmix64(kmer + seed) is a multiple of
    # s (src/lib.rs:796-806), in scan order: the sample at 8 is a subsequence of
    # the one at 4, and each keeps about 1/s of the k-mers
    seq = _sequence(20_000, 11)[:9_000] + b"N" + _sequence(20_000, 11)[9_001:]
    k, seed = 31, 42
    every = zr.scan_kmers(seq, k).tolist()
    for s in (2, 4, 8, 5):
        got = zr.read_features(seq, k, seed, mode=2, sampling=s).tolist()
        assert got == [x for x in every if sm64(x + seed) % s == 0]
        n, p = len(every), 1.0 / s
        assert abs(len(got) - n * p) < 6 * (n * p * (1 - p)) ** 0.5
    s4 = zr.read_features(seq, k, seed, mode=2, sampling=4).tolist()
    s8 = iter(s4)
    assert all(x in s8 for x in zr.read_features(seq, k, seed, mode=2, sampling=8).tolist())
    # canonical k-mers: the reverse complement samples the same multiset
    rc = zr.read_features(_rc(seq), k, seed, mode=2, sampling=4).tolist()
    assert sorted(rc) == sorted(s4)


def test_positions_closed_form():
    # storage.rs:4800-4826 (legacy segment length) plus fixed-length geometry
    for hashes in (2, 3, 4, 8, 32):
        for exponent in range(5, 25):
            cells = 1 << exponent
            for key in (0, 1, 123, MASK64):
                for fixed in (0, 64):
                    if fixed and cells // fixed < hashes:
                        continue
                    if fixed:
                        length = fixed
                    else:
                        n = min(max(cells // hashes, 1), 4096)
                        length = 1 << (n.bit_length() - 1)
                    count = cells // length - hashes + 1
                    h1 = oracle().zo_mixsplit(key, 69)
                    h2 = sm64(h1)
                    base = ((h1 * count) >> 64) * length
                    h = h1
                    expected = []
                    for i in range(hashes):
                        expected.append(base + i * length + ((h ^ (h >> 33)) & (length - 1)))
                        h = (h + h2) & MASK64
                    out = np.zeros(hashes, np.uint64)
                    oracle().zo_positions(key, 69, cells, hashes, fixed, out.ctypes.data_as(u64p))
                    assert out.tolist() == expected
                    assert all(0 <= s < cells for s in expected)
                    assert len(set(s // length for s in expected)) == hashes


def test_abandoned_hash_collision_golden():
    v = vectors()["abandoned_hash_collision"]
    a, b = v["equal"]
    h = oracle().zo_abandoned_hash
    assert h(a, v["seed"]) == h(b, v["seed"]) != h(v["different"], v["seed"])


def test_fingerprint_and_partition_closed_forms():
    for key in (0, 1, 123, MASK64):
        for bits in (1, 2, 4, 8, 16, 32):
            assert oracle().zo_fingerprint(key, 69, bits) == \
                sm64(key ^ 69 ^ 0x8E2D4F6A9B13C5D7) & ((1 << bits) - 1)
        assert oracle().zo_partition_id(key, 69, 8) == sm64(key ^ 69 ^ 0x34B61D8F29A7C053) & 7


def test_get_put_bits_round_trip():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        width = int(rng.integers(1, 65))
        offset = int(rng.integers(0, 200))
        row = C.create_string_buffer(40)
        value = int(rng.integers(0, 2**63, dtype=np.uint64)) * 2 + int(rng.integers(0, 2))
        value &= (1 << width) - 1
        oracle().zo_put_bits(row, 40, offset, width, value)
        assert oracle().zo_get_bits(row.raw, offset, width) == value
        as_int = int.from_bytes(row.raw, "little")
        assert (as_int >> offset) & ((1 << width) - 1) == value
        assert as_int & ~(((1 << width) - 1) << offset) == 0


def test_fused_screen_matches_materialized_rows():
    # storage.rs:4756-4797
    for bits in (1, 2, 4, 8, 16, 32):
        for count in (1, 7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 64, 65, 257):
            for hashes in (2, 3, 4, 8, 32):
                nbytes = (count * bits + 7) // 8
                storage = [bytes(sm64(slot * 1024 + byte) & 0xFF for byte in range(nbytes))
                           for slot in range(hashes)]
                rows = (C.c_void_p * hashes)(*[C.cast(C.c_char_p(s), C.c_void_p) for s in storage])
                keep = storage
                mat = bytearray(nbytes)
                for s in storage:
                    for i, x in enumerate(s):
                        mat[i] ^= x
                for fp in (0, 1, (1 << bits) - 1):
                    out_a = np.zeros(count, np.uint32)
                    out_b = np.zeros(count, np.uint32)
                    na = oracle().zo_screen_rows(rows, hashes, bits, fp, count,
                                                 out_a.ctypes.data_as(u32p))
                    nb = oracle().zo_screen_materialized(bytes(mat), bits, fp, count,
                                                         out_b.ctypes.data_as(u32p))
                    assert out_a[:na].tolist() == out_b[:nb].tolist()
                del keep
                zero = bytes(nbytes)
                rows = (C.c_void_p * 2)(*[C.cast(C.c_char_p(zero), C.c_void_p)] * 2)
                out = np.zeros(count, np.uint32)
                n = oracle().zo_screen_rows(rows, 2, bits, 0, count, out.ctypes.data_as(u32p))
                assert out[:n].tolist() == list(range(count))


def _union_index(key, recovered, recovered_hash_key=None):
    """storage.rs:5232-5340: a 4-slot, H=2, b=8 legacy bank with one planted
    two-member group whose XOR yields fingerprint | mask 0b01."""
    seed = 69
    data = C.create_string_buffer(8)
    slots = np.zeros(2, np.uint64)
    oracle().zo_positions(key, seed, 4, 2, 0, slots.ctypes.data_as(u64p))
    value = oracle().zo_fingerprint(key, seed, 8) | (1 << 8)
    row = C.cast(C.addressof(data) + int(slots[0]) * 2, C.c_void_p)
    oracle().zo_put_bits(row, 2, 0, 10, value)
    manifest = {
        "k": 31, "hashes": 2, "fingerprint_bits": 8, "index_mode": 0, "minimizer_size": 31,
        "seed": seed, "segment_length": None, "colors": [{"name": "a"}, {"name": "b"}],
        "banks": [{"slots": 4, "row_bits": 10, "fingerprint_first": False,
                   "groups": [{"members": [0, 1], "width": 10, "offset": 0}]}],
    }
    rec = None
    if recovered:
        hk = key if recovered_hash_key is None else recovered_hash_key
        rec = (np.array([oracle().zo_abandoned_hash(hk, seed)], np.uint32),
               np.array([0, len(recovered)], np.uint64), np.array(recovered, np.uint32))
    payload = np.frombuffer(data.raw, np.uint8).copy()
    return zr.OracleIndex(manifest, [payload], rec)


def test_union_index_recovery_known_answers():
    v = vectors()["union_index_recovery"]
    for case in v["cases"]:
        idx = _union_index(v["key"], case["recovered"])
        scores = np.zeros(2, np.uint64)
        idx.lookup(v["key"], scores)
        assert scores.tolist() == case["expected"]
        idx.lookup(v["key"], scores)
        assert scores.tolist() == [2 * x for x in case["expected"]]


def test_hash_collision_unions_origins():
    v = vectors()["collision_union_lookup"]
    idx = _union_index(v["query_key"], v["recovered"], recovered_hash_key=v["recovered_key"])
    scores = np.zeros(2, np.uint64)
    idx.lookup(v["query_key"], scores)
    assert scores.tolist() == v["expected"]


@pytest.mark.parametrize("bits", [1, 2, 4, 8, 16, 32])
def test_plane_layout_equals_legacy_layout(bits):
    # storage.rs:5343-5423
    groups, nxt, offset = [], 0, 0
    for i in range(40):
        count = i % 3 + 1
        width = bits + (0 if count == 1 else count)
        groups.append({"members": list(range(nxt, nxt + count)), "width": width, "offset": offset})
        nxt += count
        offset += width
    stride = (offset + 7) // 8
    legacy = C.create_string_buffer(4 * stride)
    planes = C.create_string_buffer(4 * stride)
    for slot in range(4):
        for i, g in enumerate(groups):
            value = sm64(slot * 40 + i) & ((1 << g["width"]) - 1)
            base_l = C.addressof(legacy) + slot * stride
            base_p = C.addressof(planes) + slot * stride
            oracle().zo_put_bits(base_l, stride, g["offset"], g["width"], value)
            oracle().zo_put_bits(base_p, stride, i * bits, bits, value & ((1 << bits) - 1))
            if g["width"] > bits:
                oracle().zo_put_bits(base_p, stride, 40 * bits + g["offset"] - i * bits,
                                     g["width"] - bits, value >> bits)
    def make(layout, payload):
        man = {"k": 31, "hashes": 2, "fingerprint_bits": bits, "index_mode": 0,
               "minimizer_size": 31, "seed": 69, "segment_length": None,
               "colors": [{"name": str(i)} for i in range(nxt)],
               "banks": [{"slots": 4, "row_bits": offset, "fingerprint_first": layout,
                          "groups": groups}]}
        rec = (np.array([oracle().zo_abandoned_hash(123, 69)], np.uint32),
               np.array([0, 1], np.uint64), np.array([1], np.uint32))
        return zr.OracleIndex(man, [np.frombuffer(payload.raw, np.uint8).copy()], rec)
    old, new = make(False, legacy), make(True, planes)
    for key in range(20000):
        a = np.zeros(nxt, np.uint64)
        b = np.zeros(nxt, np.uint64)
        old.lookup(key, a)
        new.lookup(key, b)
        assert a.tolist() == b.tolist(), key


def test_tsv_formatter_matches_reference_bytes():
    v = vectors()["tsv_record_output"]
    colors = [(v["color_name"], 0)]
    hit = v["hit_record"]
    counts = [[c for _, c in hit["hits"]]]
    text = zr.records_text([hit["name"]], [hit["total"]], counts, colors, 0.0, "tsv")
    assert text.split("\n", 1)[1] == hit["expected"]
    empty = v["empty_record"]
    text = zr.records_text([empty["name"]], [empty["total"]], [[]], [], 0.5, "tsv")
    assert text.split("\n", 1)[1] == empty["expected"]


def test_ratio_formatting_ties_round_to_even():
    # 1/512 = 0.001953125 exactly: the 9th decimal is a tie
    assert zr.fmt_ratio(1, 512) == "0.00195312"
    assert zr.fmt_ratio(3, 512) == "0.00585938"


def _random_plane_index(rng, bits, hashes, n_groups, merged_every, partitions, recovered_keys,
                        seed=77, segment_length=64):
    """A plane-layout (fingerprint_first) index with random rows: singleton
    groups and every `merged_every`-th group a 3-member group, split into
    `partitions` routed banks; the recovery map lists `recovered_keys`."""
    groups, nxt, offset = [], 0, 0
    for i in range(n_groups):
        count = 3 if merged_every and i % merged_every == merged_every - 1 else 1
        width = bits + (0 if count == 1 else count)
        groups.append({"members": list(range(nxt, nxt + count)), "width": width, "offset": offset})
        nxt += count
        offset += width
    row_bytes = (offset + 7) // 8
    slots = 64 * (16 + hashes - 1)
    banks, payloads = [], []
    for p in range(partitions):
        meta = {"slots": slots, "row_bits": offset, "fingerprint_first": True, "groups": groups}
        if partitions > 1:
            meta["routing"] = {"partition": p, "partitions": partitions}
        banks.append(meta)
        payloads.append(rng.integers(0, 256, slots * row_bytes, dtype=np.uint8))
    manifest = {"k": 31, "hashes": hashes, "fingerprint_bits": bits, "index_mode": 0,
                "minimizer_size": 31, "seed": seed, "segment_length": segment_length,
                "colors": [{"name": str(i)} for i in range(nxt)], "banks": banks}
    hs = sorted({oracle().zo_abandoned_hash(int(k), seed) for k in recovered_keys})
    origins = [sorted(rng.choice(nxt, int(rng.integers(1, 5)), replace=False).tolist()) for _ in hs]
    off = np.zeros(len(hs) + 1, np.uint64)
    off[1:] = np.cumsum([len(o) for o in origins])
    rec = (np.array(hs, np.uint32), off,
           np.array([c for o in origins for c in o] or [0], np.uint32)) if hs else None
    return zr.OracleIndex(manifest, payloads, rec)


@pytest.mark.parametrize("bits,hashes,merged_every,partitions",
                         [(8, 4, 0, 1), (8, 3, 5, 1), (16, 4, 0, 2), (32, 2, 7, 1), (4, 4, 0, 1),
                          (8, 4, 0, 3)])
def test_fast_cpu_lookup_equals_oracle(bits, hashes, merged_every, partitions):
    """The timed CPU baseline (blocks of 16 keys, rows prefetched 4 ahead,
    AVX2 plane screen) counts exactly what the scalar oracle counts, with
    duplicate keys, recovered keys, merged groups, partitions and a plane
    length that is not a multiple of 32 bytes."""
    rng = np.random.default_rng(bits * 100 + hashes * 10 + partitions)
    n_reads = 40
    lens = rng.integers(0, 300, n_reads)
    pool = rng.integers(0, 1 << 62, 500, dtype=np.uint64)  # repeats within reads
    keys = pool[rng.integers(0, len(pool), int(lens.sum()))]
    key_off = np.zeros(n_reads + 1, np.uint64)
    key_off[1:] = np.cumsum(lens)
    idx = _random_plane_index(rng, bits, hashes, 613 if bits == 8 else 301, merged_every,
                              partitions, pool[::7])
    want = np.zeros((n_reads, idx.n_colors), np.uint64)
    got = np.zeros((n_reads, idx.n_colors), np.uint64)
    oracle().zo_lookup_keys(C.byref(idx.z), keys.ctypes.data_as(u64p),
                            key_off.ctypes.data_as(u64p), n_reads, want.ctypes.data_as(u64p), 2)
    oracle().zo_lookup_keys_fast(C.byref(idx.z), keys.ctypes.data_as(u64p),
                                 key_off.ctypes.data_as(u64p), n_reads, got.ctypes.data_as(u64p), 3)
    assert want.sum() > 0
    assert np.array_equal(want, got)
