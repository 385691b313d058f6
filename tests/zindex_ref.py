"""Test-side helpers: the CPU oracle binding (oracle/liboracle.so), an
independent reader of the on-disk index format, and restatements of the
reference's output formatting. Used only by tests, smoke() and bench.py's
cpu_baseline leg as the checker.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
GOLDEN = Path(__file__).resolve().parent / "golden"

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class ZoBank(C.Structure):
    _fields_ = [
        ("data", C.c_void_p), ("slots", C.c_uint64), ("row_bits", C.c_uint64),
        ("n_groups", C.c_uint32), ("member_off", u32p), ("members", u32p), ("width", u32p),
        ("offset", u64p), ("fingerprint_first", C.c_int), ("partition", C.c_uint32),
        ("partitions", C.c_uint32),
    ]


class ZoRecovery(C.Structure):
    _fields_ = [("n", C.c_uint64), ("hashes", u32p), ("off", u64p), ("origins", u32p)]


class ZoIndex(C.Structure):
    _fields_ = [
        ("k", C.c_int), ("hashes", C.c_int), ("bits", C.c_int), ("mode", C.c_int),
        ("m", C.c_int), ("seed", C.c_uint64), ("segment_length", C.c_uint64),
        ("n_colors", C.c_uint32), ("banks", C.POINTER(ZoBank)), ("n_banks", C.c_uint32),
        ("recovery", C.POINTER(ZoRecovery)), ("sampling", C.c_uint64),
        ("hpc", C.c_int), ("seed_width", C.c_uint32), ("seed_mask", C.c_uint64),
        ("findere_z", C.c_uint32), ("strobe_min", C.c_uint64), ("strobe_max", C.c_uint64),
    ]


HITS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64))


def spaced_seed(pattern):
    """(width, mask) of a '#'/'_' pattern: two set bits per retained base, the
    first base most significant (src/anchors.rs:15-43); (0, 0) for None."""
    if not pattern:
        return 0, 0
    mask = 0
    for ch in pattern:
        mask = (mask << 2) | (3 if ch == "#" else 0)
    return len(pattern), mask


_oracle = None


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return ORACLE_LIB


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            build_oracle()
        o = C.CDLL(str(ORACLE_LIB))
        o.zo_splitmix64.restype = C.c_uint64
        o.zo_splitmix64.argtypes = [C.c_uint64]
        o.zo_mixsplit.restype = C.c_uint64
        o.zo_mixsplit.argtypes = [C.c_uint64, C.c_uint64]
        o.zo_encode_base.restype = C.c_int
        o.zo_encode_base.argtypes = [C.c_uint8]
        o.zo_encode_full_kmer.restype = C.c_int
        o.zo_encode_full_kmer.argtypes = [C.c_char_p, C.c_size_t, u64p]
        o.zo_canonical_encoded.restype = C.c_uint64
        o.zo_canonical_encoded.argtypes = [C.c_uint64, C.c_int]
        o.zo_fingerprint.restype = C.c_uint64
        o.zo_fingerprint.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
        o.zo_abandoned_hash.restype = C.c_uint32
        o.zo_abandoned_hash.argtypes = [C.c_uint64, C.c_uint64]
        o.zo_partition_id.restype = C.c_uint64
        o.zo_partition_id.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        o.zo_positions.restype = None
        o.zo_positions.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_uint64, u64p]
        o.zo_get_bits.restype = C.c_uint64
        o.zo_get_bits.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t]
        o.zo_put_bits.restype = None
        o.zo_put_bits.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint64]
        o.zo_scan_kmers.restype = C.c_size_t
        o.zo_scan_kmers.argtypes = [C.c_char_p, C.c_size_t, C.c_int, u64p]
        o.zo_scan_ocm.restype = C.c_size_t
        o.zo_scan_ocm.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.c_int, C.c_uint64, u64p, u64p]
        o.zo_scan_ocm_bruteforce.restype = C.c_size_t
        o.zo_scan_ocm_bruteforce.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.c_int, C.c_uint64,
                                             u64p, u64p]
        o.zo_homopolymer_compress.restype = C.c_size_t
        o.zo_homopolymer_compress.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
        o.zo_spaced_project.restype = C.c_uint64
        o.zo_spaced_project.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        o.zo_spaced_canonical.restype = C.c_uint64
        o.zo_spaced_canonical.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        o.zo_strobe_hash.restype = C.c_uint64
        o.zo_strobe_hash.argtypes = [C.c_uint64]
        o.zo_link_strobes.restype = None
        o.zo_link_strobes.argtypes = [u64p, u64p, C.c_size_t, C.c_uint64, C.c_uint64, u64p]
        o.zo_findere_sequence.restype = None
        o.zo_findere_sequence.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.c_int, C.c_uint32,
                                          HITS_FN, C.c_void_p, u64p, u64p, u64p]
        o.zo_read_features.restype = C.c_size_t
        o.zo_read_features.argtypes = [C.POINTER(ZoIndex), C.c_char_p, C.c_size_t, u64p]
        o.zo_lookup.restype = None
        o.zo_lookup.argtypes = [C.POINTER(ZoIndex), C.c_uint64, C.c_uint64, u64p]
        o.zo_query_reads.restype = C.c_int
        o.zo_query_reads.argtypes = [C.POINTER(ZoIndex), C.c_void_p, u64p, C.c_uint64, u64p, u64p,
                                     C.c_int]
        o.zo_query_aggregate.restype = C.c_uint64
        o.zo_query_aggregate.argtypes = [C.POINTER(ZoIndex), C.c_void_p, u64p, C.c_uint64, u64p,
                                         C.c_int]
        o.zo_lookup_keys.restype = C.c_int
        o.zo_lookup_keys.argtypes = [C.POINTER(ZoIndex), u64p, u64p, C.c_uint64, u64p, C.c_int]
        o.zo_lookup_keys_fast.restype = C.c_int
        o.zo_lookup_keys_fast.argtypes = [C.POINTER(ZoIndex), u64p, u64p, C.c_uint64, u64p,
                                          C.c_int]
        o.zo_screen_rows.restype = C.c_size_t
        o.zo_screen_rows.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_uint64,
                                     C.c_size_t, u32p]
        o.zo_screen_materialized.restype = C.c_size_t
        o.zo_screen_materialized.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.c_size_t, u32p]
        _oracle = o
    return _oracle


def scan_kmers(seq: bytes, k: int) -> np.ndarray:
    out = np.zeros(max(len(seq), 1), np.uint64)
    n = oracle().zo_scan_kmers(seq, len(seq), k, out.ctypes.data_as(u64p))
    return out[:n]


def findere_sequence(seq: bytes, k: int, z: int, n_colors: int, hits_fn):
    """findere::Scorer::sequence with hits_fn(key) -> list of per-colour hits;
    returns (scores, total_kmers, smer_queries)."""
    def cb(_ctx, key, hits):
        for c, v in enumerate(hits_fn(key)):
            hits[c] = v
    fn = HITS_FN(cb)
    scores = np.zeros(max(n_colors, 1), np.uint64)
    total, probes = C.c_uint64(0), C.c_uint64(0)
    oracle().zo_findere_sequence(seq, len(seq), k, z, n_colors, fn, None,
                                 scores.ctypes.data_as(u64p), C.byref(total), C.byref(probes))
    return scores[:n_colors].tolist(), total.value, probes.value


def homopolymer_compress(seq: bytes) -> bytes:
    out = C.create_string_buffer(max(len(seq), 1))
    n = oracle().zo_homopolymer_compress(seq, len(seq), out)
    return out.raw[:n]


def link_strobes(pos, hashes, lo: int, hi: int) -> list:
    pos = np.ascontiguousarray(pos, np.uint64)
    hashes = np.ascontiguousarray(hashes, np.uint64)
    keys = np.zeros(max(len(pos), 1), np.uint64)
    oracle().zo_link_strobes(pos.ctypes.data_as(u64p), hashes.ctypes.data_as(u64p), len(pos),
                             lo, hi, keys.ctypes.data_as(u64p))
    return keys[:len(pos)].tolist()


def read_features(seq: bytes, k: int, seed: int, mode: int = 0, m: int = 0,
                  sampling: int = 16, hpc: bool = False, pattern=None,
                  strobes=None) -> np.ndarray:
    """The query-side features of one read (k-mers, OCM, modimizers, with an
    optional pattern; strobes=(min, max) links randstrobemers)."""
    width, mask = spaced_seed(pattern)
    lo, hi = strobes or (0, 0)
    idx = ZoIndex(k=k, mode=mode, m=m, seed=seed, sampling=sampling, hpc=int(hpc),
                  seed_width=width, seed_mask=mask, strobe_min=lo, strobe_max=hi)
    out = np.zeros(max(len(seq), 1), np.uint64)
    n = oracle().zo_read_features(C.byref(idx), seq, len(seq), out.ctypes.data_as(u64p))
    return out[:n]


def scan_ocm(seq: bytes, k: int, m: int, seed: int):
    pos = np.zeros(max(len(seq), 1), np.uint64)
    key = np.zeros(max(len(seq), 1), np.uint64)
    n = oracle().zo_scan_ocm(seq, len(seq), k, m, seed, pos.ctypes.data_as(u64p),
                             key.ctypes.data_as(u64p))
    return pos[:n], key[:n]


def scan_ocm_bruteforce(run: bytes, k: int, m: int, seed: int):
    pos = np.zeros(max(len(run), 1), np.uint64)
    key = np.zeros(max(len(run), 1), np.uint64)
    n = oracle().zo_scan_ocm_bruteforce(run, len(run), k, m, seed, pos.ctypes.data_as(u64p),
                                        key.ctypes.data_as(u64p))
    return pos[:n], key[:n]


# ---------------------------------------------------------------------------
# zstd (ctypes over libzstd.so.1) and an independent index reader
# ---------------------------------------------------------------------------
class _InBuf(C.Structure):
    _fields_ = [("src", C.c_void_p), ("size", C.c_size_t), ("pos", C.c_size_t)]


class _OutBuf(C.Structure):
    _fields_ = [("dst", C.c_void_p), ("size", C.c_size_t), ("pos", C.c_size_t)]


_zstd = None


def _libzstd():
    global _zstd
    if _zstd is None:
        z = C.CDLL("libzstd.so.1")
        z.ZSTD_createDCtx.restype = C.c_void_p
        z.ZSTD_freeDCtx.argtypes = [C.c_void_p]
        z.ZSTD_decompressStream.restype = C.c_size_t
        z.ZSTD_decompressStream.argtypes = [C.c_void_p, C.POINTER(_OutBuf), C.POINTER(_InBuf)]
        z.ZSTD_isError.restype = C.c_uint
        z.ZSTD_isError.argtypes = [C.c_size_t]
        z.ZSTD_compress.restype = C.c_size_t
        z.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
        z.ZSTD_compressBound.restype = C.c_size_t
        z.ZSTD_compressBound.argtypes = [C.c_size_t]
        _zstd = z
    return _zstd


def zstd_decompress(data: bytes) -> bytes:
    z = _libzstd()
    d = z.ZSTD_createDCtx()
    src = C.create_string_buffer(data, len(data))
    inb = _InBuf(C.cast(src, C.c_void_p), len(data), 0)
    chunk = C.create_string_buffer(1 << 20)
    out = bytearray()
    try:
        while True:
            ob = _OutBuf(C.cast(chunk, C.c_void_p), len(chunk), 0)
            r = z.ZSTD_decompressStream(d, C.byref(ob), C.byref(inb))
            if z.ZSTD_isError(r):
                raise ValueError("zstd decode error")
            out += chunk.raw[: ob.pos]
            if inb.pos >= inb.size and ob.pos < len(chunk):
                break
    finally:
        z.ZSTD_freeDCtx(d)
    return bytes(out)


def zstd_compress(data: bytes, level: int = 3) -> bytes:
    z = _libzstd()
    cap = z.ZSTD_compressBound(len(data))
    dst = C.create_string_buffer(cap)
    src = C.create_string_buffer(data, len(data))
    n = z.ZSTD_compress(dst, cap, src, len(data), level)
    assert not z.ZSTD_isError(n)
    return dst.raw[:n]


def to_v1(src_dir, dst_dir, fallback) -> None:
    """Rewrite an index directory (built without a fixed segment length) as
    manifest version 1: legacy interleaved banks whose streams end with every
    group's fallback list (storage.rs:2555-2612), no separate recovery file.
    fallback: {(bank, group): [(key, member mask), ...]} with the caller's keys."""
    to_legacy_layout(src_dir, dst_dir, fallback=fallback)


def to_legacy_layout(src_dir, dst_dir, order=None, fallback=None) -> None:
    """Rewrite an index directory with legacy interleaved banks
    (fingerprint_first = false; storage.rs:2864-2914): group g of a row holds
    fp | mask << b at bits [offset_g, offset_g + width_g), offsets cumulative
    in manifest order. `order(groups) -> permutation` reorders each bank's
    groups (default: as they are). Bank files are one zstd frame without a
    seek table, behind the legacy magic ZORPOW1 (the old single-stream
    container, bank_io.rs:351-366)."""
    import shutil
    src_dir, dst_dir = Path(src_dir), Path(dst_dir)
    dst_dir.mkdir(parents=True)
    disk = read_index_dir(src_dir)
    man = json.loads(json.dumps(disk["manifest"]))
    b = man["fingerprint_bits"]
    for bi, bank in enumerate(man["banks"]):
        groups = bank["groups"]
        G = len(groups)
        rb = (bank["row_bits"] + 7) // 8
        payload = disk["banks"][bi]
        # source plane layout: fp at g*b, masks after G*b in group order
        mw = [len(g["members"]) if len(g["members"]) > 1 else 0 for g in groups]
        mask_at, acc = [], G * b
        for w in mw:
            mask_at.append(acc)
            acc += w
        perm = list(range(G)) if order is None else list(order(groups))
        new_groups, offsets, off = [], [], 0
        for g in perm:
            ng = dict(groups[g])
            ng["offset"] = off
            offsets.append(off)
            off += ng["width"]
            new_groups.append(ng)
        out = bytearray(len(payload))
        fpm = (1 << b) - 1
        for r in range(bank["slots"]):
            row = int.from_bytes(payload[r * rb:(r + 1) * rb].tobytes(), "little")
            v = 0
            for g, o in zip(perm, offsets):
                fp = (row >> (g * b)) & fpm
                mask = (row >> mask_at[g]) & ((1 << mw[g]) - 1) if mw[g] else 0
                v |= (fp | (mask << b)) << o
            out[r * rb:(r + 1) * rb] = v.to_bytes(rb, "little")
        bank["groups"] = new_groups
        bank["fingerprint_first"] = False
        header = b"ZORPOW1\0" + np.array([bank["slots"], bank["row_bits"], G, len(out)],
                                         dtype="<u8").tobytes()
        trailer = b""
        if fallback is not None:
            for gi, g in enumerate(new_groups):
                entries = sorted(fallback.get((bi, gi), []))
                g.pop("abandoned_keys", None)
                g["fallback_keys"] = len(entries)
                trailer += np.array([len(entries)], dtype="<u8").tobytes()
                for key, mask in entries:
                    value = oracle().zo_fingerprint(int(key), man["seed"], b) | (
                        0 if len(g["members"]) == 1 else mask << b)
                    trailer += np.array([key, value], dtype="<u8").tobytes()
        (dst_dir / bank["file"]).write_bytes(zstd_compress(header + bytes(out) + trailer))
    if fallback is not None:
        assert man.get("segment_length") is None
        man["version"] = 1
        for f in ("abandoned_recovery", "abandoned_file", "abandoned_hashes",
                  "abandoned_origins"):  # newer fields; recovery defaults on
            man.pop(f, None)
    elif man.get("abandoned_file"):
        shutil.copy(src_dir / man["abandoned_file"], dst_dir / man["abandoned_file"])
    (dst_dir / "manifest.json").write_text(json.dumps(man, indent=2))


def read_varint(buf: bytes, at: int):
    value = 0
    for shift in range(0, 35, 7):
        b = buf[at]
        at += 1
        value |= (b & 127) << shift
        if not b & 128:
            return value, at
    raise ValueError("varint overflow")


def read_index_dir(path) -> dict:
    """manifest + decoded bank payloads + recovery map, read independently of
    the product (bank_io.rs framing is transparent to a streaming decoder)."""
    path = Path(path)
    man = json.loads((path / "manifest.json").read_text())
    banks = []
    records = []  # version 1: (abandoned hash, origins) of every fallback key
    for b in man["banks"]:
        raw = zstd_decompress((path / b["file"]).read_bytes())
        magic = raw[:8]
        assert magic in (b"ZORPOW3\0", b"ZORPOW1\0"), magic
        slots, row_bits, groups, n = np.frombuffer(raw[8:40], dtype="<u8")
        assert (slots, row_bits, groups) == (b["slots"], b["row_bits"], len(b["groups"]))
        assert n == slots * ((row_bits + 7) // 8)
        payload = np.frombuffer(raw[40:40 + int(n)], dtype=np.uint8)
        assert len(payload) == n
        tail = raw[40 + int(n):]
        if man["version"] == 1:
            # per group: u64 count, then count x (u64 key, u64 value) (storage.rs:2564-2612)
            at, fb = 0, man["fingerprint_bits"]
            for g in b["groups"]:
                (cnt,) = np.frombuffer(tail[at:at + 8], dtype="<u8")
                at += 8
                assert cnt == g["fallback_keys"]
                kv = np.frombuffer(tail[at:at + 16 * int(cnt)], dtype="<u8").reshape(-1, 2)
                at += 16 * int(cnt)
                for key, value in kv:
                    mask = 1 if len(g["members"]) == 1 else int(value) >> fb
                    origins = [g["members"][i] for i in range(64) if mask >> i & 1]
                    records.append((oracle().zo_abandoned_hash(int(key), man["seed"]), origins))
            assert at == len(tail)
        else:
            assert len(tail) == 0
        banks.append(payload)
    rec = None
    if man["version"] == 1 and man.get("abandoned_recovery", True):
        by = {}
        for h, o in records:
            by.setdefault(h, set()).update(o)
        hashes = sorted(by)
        off, origins = [0], []
        for h in hashes:
            origins += sorted(by[h])
            off.append(len(origins))
        rec = (np.array(hashes, np.uint32), np.array(off, np.uint64), np.array(origins, np.uint32))
    if man.get("abandoned_recovery", True) and man.get("abandoned_file"):
        raw = zstd_decompress((path / man["abandoned_file"]).read_bytes())
        assert raw[:8] == b"ZORABN2\0"
        at, prev = 8, 0
        hashes, off, origins = [], [0], []
        for _ in range(man["abandoned_hashes"]):
            d, at = read_varint(raw, at)
            prev += d
            cnt, at = read_varint(raw, at)
            o = 0
            for _ in range(cnt):
                d, at = read_varint(raw, at)
                o += d
                origins.append(o)
            hashes.append(prev)
            off.append(len(origins))
        assert at == len(raw)
        rec = (np.array(hashes, np.uint32), np.array(off, np.uint64), np.array(origins, np.uint32))
    return {"manifest": man, "banks": banks, "recovery": rec}


class OracleIndex:
    """The oracle's view of an index (manifest + payloads + recovery)."""

    def __init__(self, manifest: dict, banks: list, recovery=None, recover: bool = True):
        self.manifest = manifest
        self._keep = []
        zb = (ZoBank * len(manifest["banks"]))()
        for i, (meta, payload) in enumerate(zip(manifest["banks"], banks)):
            payload = np.ascontiguousarray(payload, dtype=np.uint8)
            groups = meta["groups"]
            member_off = np.zeros(len(groups) + 1, np.uint32)
            member_off[1:] = np.cumsum([len(g["members"]) for g in groups])
            members = np.array([m for g in groups for m in g["members"]] or [0], np.uint32)
            width = np.array([g["width"] for g in groups], np.uint32)
            offset = np.array([g["offset"] for g in groups], np.uint64)
            self._keep += [payload, member_off, members, width, offset]
            routing = meta.get("routing")
            zb[i] = ZoBank(payload.ctypes.data, meta["slots"], meta["row_bits"], len(groups),
                           member_off.ctypes.data_as(u32p), members.ctypes.data_as(u32p),
                           width.ctypes.data_as(u32p), offset.ctypes.data_as(u64p),
                           int(meta.get("fingerprint_first", False)),
                           routing["partition"] if routing else 0,
                           routing["partitions"] if routing else 1)
        self._banks = zb
        self._rec = None
        if recover and recovery is not None and len(recovery[0]):
            h, off, o = (np.ascontiguousarray(x) for x in recovery)
            self._keep += [h, off, o]
            self._rec = ZoRecovery(len(h), h.ctypes.data_as(u32p), off.ctypes.data_as(u64p),
                                   o.ctypes.data_as(u32p))
        mode = manifest["index_mode"]
        self.z = ZoIndex(manifest["k"], manifest["hashes"], manifest["fingerprint_bits"], mode,
                         manifest["minimizer_size"], manifest["seed"],
                         manifest.get("segment_length") or 0, len(manifest["colors"]), zb,
                         len(manifest["banks"]),
                         C.pointer(self._rec) if self._rec is not None else None,
                         manifest.get("modimizer_sampling") or 16,
                         int(bool(manifest.get("homopolymer_compression"))),
                         *spaced_seed(manifest.get("seed_pattern")),
                         manifest.get("findere_z", 10),
                         *((manifest["randstrobemers"]["min_distance"],
                            manifest["randstrobemers"]["max_distance"])
                           if manifest.get("randstrobemers") else (0, 0)))

    @classmethod
    def from_dir(cls, path, recover: bool = True) -> "OracleIndex":
        d = read_index_dir(path)
        return cls(d["manifest"], d["banks"], d["recovery"], recover)

    @property
    def n_colors(self) -> int:
        return len(self.manifest["colors"])

    def query_reads(self, buf: np.ndarray, off: np.ndarray, threads: int = 1):
        n = len(off) - 1
        counts = np.zeros((max(n, 1), self.n_colors), np.uint64)
        totals = np.zeros(max(n, 1), np.uint64)
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        off = np.ascontiguousarray(off - off[0], dtype=np.uint64) if n else off
        oracle().zo_query_reads(C.byref(self.z), buf.ctypes.data + int(0), off.ctypes.data_as(u64p),
                                n, counts.ctypes.data_as(u64p), totals.ctypes.data_as(u64p),
                                threads)
        return totals[:n], counts[:n]

    def lookup(self, key: int, scores: np.ndarray, weight: int = 1) -> None:
        oracle().zo_lookup(C.byref(self.z), key, weight, scores.ctypes.data_as(u64p))


# ---------------------------------------------------------------------------
# Output restatements (storage.rs:3817-3844, record_output.rs:14-84)
# ---------------------------------------------------------------------------
def fmt_ratio(count: int, total: int) -> str:
    return "%.8f" % (0.0 if total == 0 else count / total)


def tsv_field(value: str) -> str:
    if any(c in value for c in "\t\r\n\""):
        return '"' + value.replace('"', '""') + '"'
    return value


def aggregate_text(colors: list, agg, total: int, min_ratio: float) -> str:
    lines = ["color\tmatches\tratio\tunique_kmers"]
    for (name, unique), m in zip(colors, agg):
        ratio = 0.0 if total == 0 else int(m) / total
        if ratio >= min_ratio:
            lines.append(f"{name}\t{int(m)}\t{fmt_ratio(int(m), total)}\t{unique}")
    return "\n".join(lines) + "\n"


def records_text(names, totals, counts, colors, min_ratio: float, fmt: str) -> str:
    out = ["query_id\tdataset_id\tdataset_name\tmatches\tqueried_seeds\tratio" if fmt == "tsv"
           else "query_id\tvalid_kmers\thits"]
    for name, total, row in zip(names, totals, counts):
        total = int(total)
        hits = [(c, int(v)) for c, v in enumerate(row)
                if min_ratio == 0.0 or (total != 0 and int(v) / total >= min_ratio)]
        if fmt == "compact":
            out.append(f"{name}\t{total}\t" + ",".join(f"{c}:{v}" for c, v in hits))
        elif hits:
            for c, v in hits:
                out.append(f"{tsv_field(name)}\t{c}\t{tsv_field(colors[c][0])}\t{v}\t{total}\t"
                           f"{fmt_ratio(v, total)}")
        else:
            out.append(f"{tsv_field(name)}\t.\t.\t0\t{total}\t0.00000000")
    return "\n".join(out) + "\n"


def read_fasta_records(path) -> list[tuple[str, bytes]]:
    """Minimal FASTA reader for fixtures (id = first token)."""
    import gzip
    p = Path(path)
    data = gzip.decompress(p.read_bytes()) if p.suffix == ".gz" else p.read_bytes()
    recs, name, seq = [], None, []
    for line in data.split(b"\n"):
        line = line.rstrip(b"\r")
        if line.startswith(b">"):
            if name is not None:
                recs.append((name, b"".join(seq)))
            name, seq = line[1:].split()[0].decode(), []
        elif line:
            seq.append(line.strip())
    if name is not None:
        recs.append((name, b"".join(seq)))
    return recs


def vectors() -> dict:
    return json.loads((GOLDEN / "reference_vectors.json").read_text())


def parse_int(v) -> int:
    return int(v, 16) if isinstance(v, str) else int(v)


def cpu_count() -> int:
    """Host threads for the oracle: affinity count capped at OMP_NUM_THREADS
    (default 16) so a shared host is not oversubscribed."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(n, cap))
