"""zindex_amd -- MI355X (gfx950) query path for zorindex indexes.

The Python surface mirrors the reference crate's public functions
(src/power_merge/mod.rs:45-49): query_command, query_command_with_memory,
query_records_command, query_records_command_with_memory, info_command and
build, with the same argument meaning and error behaviour (invalid input fails
before any output). Everything runs through the C ABI of libzindex_amd.so
(include/zindex_gpu.h); queries execute as HIP kernels on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np

from ._lib import (BatchStats, BuildConfig, IndexMeta, RequestStats, ServerConfig, ServerStats,
                   Tuning, ZG_QUERY_NO_RECOVERY, ZG_QUERY_RECORDS, ZindexError, check, lib)

__all__ = [
    "Index", "Builder", "Batch", "Server", "HostBuffer", "ZindexError", "build", "build_config",
    "query_command", "query_command_with_memory", "query_command_with_recovery",
    "query_records_command", "query_records_command_with_memory", "info_command",
    "pack_reads", "fuse_indexes", "fuse_many_indexes",
]

QUERY_MEMORY = ("all", "subindex")  # ZG_MEMORY_ALL, ZG_MEMORY_SUBINDEX
RECORD_FORMATS = {"tsv": 0, "compact": 1}


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def pack_reads(reads: Sequence[bytes] | Iterable[bytes]) -> tuple[np.ndarray, np.ndarray]:
    """Concatenate byte strings into (uint8 buffer, uint64 offsets[n+1])."""
    reads = [r if isinstance(r, (bytes, bytearray)) else bytes(r) for r in reads]
    off = np.zeros(len(reads) + 1, dtype=np.uint64)
    if reads:
        off[1:] = np.cumsum([len(r) for r in reads], dtype=np.uint64)
    buf = np.frombuffer(b"".join(reads), dtype=np.uint8) if reads else np.zeros(0, np.uint8)
    return np.ascontiguousarray(buf), off


def _as_reads(reads, offsets=None):
    if offsets is None:
        return pack_reads(reads)
    buf = np.ascontiguousarray(np.asarray(reads, dtype=np.uint8))
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
    return buf, off


def build_config(**kw) -> BuildConfig:
    """Reference CLI defaults (src/bin/zorindex.rs:55-159) with overrides."""
    cfg = BuildConfig()
    lib().zg_build_config_default(C.byref(cfg))
    for key, value in kw.items():
        if key == "index_mode" and isinstance(value, str):
            value = {"kmers": 0, "minimizers": 1, "modimizers": 2, "findere": 3}[value]
        if key == "seed_pattern" and isinstance(value, str):
            value = value.encode()
        if not hasattr(cfg, key):
            raise TypeError(f"unknown build option {key}")
        setattr(cfg, key, value)
    return cfg


class Index:
    """An index handle: manifest metadata plus (after load) HBM-resident banks."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def open(cls, directory: str) -> "Index":
        h = C.c_void_p()
        check(lib().zg_index_open(str(directory).encode(), C.byref(h)))
        return cls(h)

    def close(self) -> None:
        if self._h:
            lib().zg_index_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def meta(self) -> dict:
        m = IndexMeta()
        check(lib().zg_index_get_meta(self._h, C.byref(m)))
        return {name: getattr(m, name) for name, _ in IndexMeta._fields_}

    def colors(self) -> list[tuple[str, int]]:
        out = []
        for c in range(self.meta["n_colors"]):
            name = C.c_char_p()
            uk = C.c_uint64()
            check(lib().zg_index_color(self._h, c, C.byref(name), C.byref(uk)))
            out.append((name.value.decode(), uk.value))
        return out

    def load(self, recover: bool = True, hbm_budget: int = 0, memory: str = "all") -> "Index":
        """Makes the index resident. memory="subindex" with a budget below the
        index streams the banks that do not fit through HBM per query."""
        if memory not in QUERY_MEMORY:
            raise ValueError(f"invalid query memory mode {memory}")
        check(lib().zg_index_load_ex(self._h, int(bool(recover)), int(hbm_budget),
                                     QUERY_MEMORY.index(memory)))
        return self

    def bank_payload(self, bank: int, copy: bool = True) -> np.ndarray:
        """Decoded payload of one bank; copy=False returns a view into the
        library's host copy (valid while the index is open)."""
        p = C.POINTER(C.c_uint8)()
        n = C.c_uint64()
        check(lib().zg_index_bank_payload(self._h, bank, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, np.uint8)
        view = np.ctypeslib.as_array(p, shape=(n.value,))
        return view.copy() if copy else view

    def manifest(self) -> dict:
        import json
        text = C.c_char_p()
        n = C.c_uint64()
        check(lib().zg_index_manifest(self._h, C.byref(text), C.byref(n)))
        return json.loads(C.string_at(text, n.value).decode())

    def recovery(self):
        """(hashes, offsets, origins) of the host recovery map, or None."""
        h = C.POINTER(C.c_uint32)()
        off = C.POINTER(C.c_uint64)()
        o = C.POINTER(C.c_uint32)()
        n = C.c_uint64()
        if lib().zg_index_recovery(self._h, C.byref(h), C.byref(off), C.byref(o), C.byref(n)):
            return None
        if n.value == 0:
            return (np.zeros(0, np.uint32), np.zeros(1, np.uint64), np.zeros(0, np.uint32))
        hashes = np.ctypeslib.as_array(h, shape=(n.value,)).copy()
        offs = np.ctypeslib.as_array(off, shape=(n.value + 1,)).copy()
        origins = np.ctypeslib.as_array(o, shape=(int(offs[-1]),)).copy() if offs[-1] else \
            np.zeros(0, np.uint32)
        return hashes, offs, origins

    def write(self, out_dir: str) -> None:
        check(lib().zg_index_write(self._h, str(out_dir).encode()))

    def query_reads(self, reads, offsets=None, records: bool = False, recover: bool = True):
        """Per-read totals, per-read counts (records) and aggregate counts."""
        buf, off = _as_reads(reads, offsets)
        n = len(off) - 1
        colors = self.meta["n_colors"]
        flags = (ZG_QUERY_RECORDS if records else 0) | (0 if recover else ZG_QUERY_NO_RECOVERY)
        totals = np.zeros(max(n, 1), np.uint64)
        counts = np.zeros((max(n, 1), colors), np.uint32) if records else None
        agg = np.zeros(colors, np.uint64)
        check(lib().zg_query_reads(self._h, buf.ctypes.data_as(C.c_void_p), _u64p(off), n, flags,
                                   _u64p(totals), _u32p(counts) if records else None, _u64p(agg)))
        return totals[:n], (counts[:n] if records else None), agg

    def query_reads_noalloc(self, reads, offsets=None, records: bool = False,
                            recover: bool = True):
        """query_reads through the allocation-free entry point (a 64 MiB static
        device array per GPU; larger queries raise)."""
        buf, off = _as_reads(reads, offsets)
        n = len(off) - 1
        colors = self.meta["n_colors"]
        flags = (ZG_QUERY_RECORDS if records else 0) | (0 if recover else ZG_QUERY_NO_RECOVERY)
        totals = np.zeros(max(n, 1), np.uint64)
        counts = np.zeros((max(n, 1), colors), np.uint32) if records else None
        agg = np.zeros(colors, np.uint64)
        check(lib().zg_query_reads_noalloc(self._h, buf.ctypes.data_as(C.c_void_p), _u64p(off), n,
                                           flags, _u64p(totals),
                                           _u32p(counts) if records else None, _u64p(agg)))
        return totals[:n], (counts[:n] if records else None), agg

    def extract_seeds(self, reads, offsets=None):
        """GPU-selected seeds: (key_off[n+1], keys, read-relative positions)."""
        buf, off = _as_reads(reads, offsets)
        n = len(off) - 1
        key_off = np.zeros(n + 1, np.uint64)
        cap = int(off[-1] - off[0]) + 1 if n else 1
        keys = np.zeros(cap, np.uint64)
        pos = np.zeros(cap, np.uint64)
        check(lib().zg_extract_seeds(self._h, buf.ctypes.data_as(C.c_void_p), _u64p(off), n,
                                     _u64p(key_off), _u64p(keys), _u64p(pos), cap))
        total = int(key_off[-1])
        return key_off, keys[:total], pos[:total]

    def batch(self, reads, offsets=None, records: bool = False, recover: bool = True) -> "Batch":
        buf, off = _as_reads(reads, offsets)
        flags = (ZG_QUERY_RECORDS if records else 0) | (0 if recover else ZG_QUERY_NO_RECOVERY)
        h = C.c_void_p()
        check(lib().zg_batch_create(self._h, buf.ctypes.data_as(C.c_void_p), _u64p(off),
                                    len(off) - 1, flags, C.byref(h)))
        return Batch(self, h, len(off) - 1, records)

    def gather_probe(self, row_bytes: int, n_rows: int, bank: int = 0) -> float:
        """GB/s of random whole-row reads of row_bytes over a resident bank."""
        gbs = C.c_double()
        check(lib().zg_gather_probe(self._h, bank, row_bytes, n_rows, C.byref(gbs)))
        return gbs.value

    def gather_probe_keys(self, n_keys: int, bank: int = 0) -> float:
        """GB/s of the H rows of random keys under the bank's PositionHash geometry."""
        gbs = C.c_double()
        check(lib().zg_gather_probe_keys(self._h, bank, n_keys, C.byref(gbs)))
        return gbs.value

    @property
    def tuning(self) -> dict:
        t = Tuning()
        check(lib().zg_index_get_tuning(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    @tuning.setter
    def tuning(self, value: dict) -> None:
        t = Tuning(**value)
        check(lib().zg_index_set_tuning(self._h, C.byref(t)))


class Batch:
    """Reads resident in HBM with their seeds; run() is the persistent lookup."""

    def __init__(self, index: Index, handle: C.c_void_p, n_reads: int, records: bool):
        self.index = index
        self._h = handle
        self.n_reads = n_reads
        self.records = records

    def close(self) -> None:
        if self._h:
            lib().zg_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self) -> None:
        check(lib().zg_batch_run(self.index._h, self._h))

    def run_plain(self, first: int = 0, count: int | None = None) -> None:
        count = self.n_reads - first if count is None else count
        check(lib().zg_batch_run_plain(self.index._h, self._h, first, count))

    def reseed(self) -> None:
        check(lib().zg_batch_reseed(self.index._h, self._h))

    def autotune(self) -> dict:
        t = Tuning()
        check(lib().zg_autotune(self.index._h, self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    def results(self):
        colors = self.index.meta["n_colors"]
        n = self.n_reads
        totals = np.zeros(max(n, 1), np.uint64)
        counts = np.zeros((max(n, 1), colors), np.uint32) if self.records else None
        agg = np.zeros(colors, np.uint64)
        check(lib().zg_batch_results(self._h, _u64p(totals),
                                     _u32p(counts) if self.records else None, _u64p(agg)))
        return totals[:n], (counts[:n] if self.records else None), agg

    def store_probe(self) -> float:
        """GB/s of the last sweep chunk's hit-row stores replayed alone
        (zg_batch_store_probe); overwrites the hit rows, not the results."""
        gbs = C.c_double()
        check(lib().zg_batch_store_probe(self._h, C.byref(gbs)))
        return gbs.value

    def stats(self) -> dict:
        s = BatchStats()
        check(lib().zg_batch_get_stats(self._h, C.byref(s)))
        out = {name: getattr(s, name) for name, _ in BatchStats._fields_ if name != "tuning"}
        out["tuning"] = {name: getattr(s.tuning, name) for name, _ in Tuning._fields_}
        return out


class HostBuffer:
    """Page-locked host memory (zg_host_alloc) viewed as a numpy array."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        self._p = C.c_void_p()
        check(lib().zg_host_alloc(max(n, 1), C.byref(self._p)))
        raw = (C.c_uint8 * max(n, 1)).from_address(self._p.value)
        self.array = np.frombuffer(raw, dtype=np.uint8, count=n).view(dtype).reshape(shape)

    def free(self) -> None:
        if self._p:
            self.array = None
            lib().zg_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching request server over a resident index (zg_server_*).

    submit() queues a request and returns its ticket at once; wait(ticket)
    blocks until its results are written and returns (totals, counts,
    aggregate, stats). Buffers of queued requests are kept alive here."""

    def __init__(self, index: Index, slots: int = 2, max_requests: int = 256,
                 slot_reads: int = 32768, slot_bases: int = 64 << 20, hbm_budget: int = 0,
                 recover: bool = True):
        cfg = ServerConfig()
        lib().zg_server_config_default(C.byref(cfg))
        cfg.slots, cfg.max_requests = slots, max_requests
        cfg.slot_reads, cfg.slot_bases, cfg.hbm_budget = slot_reads, slot_bases, hbm_budget
        cfg.flags = 0 if recover else ZG_QUERY_NO_RECOVERY
        self.index = index
        self.colors = index.meta["n_colors"]
        self._h = C.c_void_p()
        self._live = {}
        check(lib().zg_server_create(index._h, C.byref(cfg), C.byref(self._h)))

    def submit(self, reads, offsets=None, records: bool = True, aggregate: bool = True,
               out=None) -> int:
        """out: optional (totals, counts) arrays to write into (e.g. HostBuffer
        views); otherwise fresh arrays are allocated."""
        buf, off = _as_reads(reads, offsets)
        n = len(off) - 1
        if out is not None:
            totals, counts = out
        else:
            totals = np.zeros(max(n, 1), np.uint64)
            counts = np.zeros((max(n, 1), self.colors), np.uint32) if records else None
        agg = np.zeros(self.colors, np.uint64) if aggregate else None
        ticket = C.c_uint64()
        check(lib().zg_server_submit(self._h, buf.ctypes.data_as(C.c_void_p), _u64p(off), n,
                                     _u64p(totals), _u32p(counts) if counts is not None else None,
                                     _u64p(agg) if agg is not None else None, C.byref(ticket)))
        self._live[ticket.value] = (buf, off, n, totals, counts, agg)
        return ticket.value

    def wait(self, ticket: int):
        st = RequestStats()
        rc = lib().zg_server_wait(self._h, ticket, C.byref(st))
        buf, off, n, totals, counts, agg = self._live.pop(ticket)
        check(rc)
        stats = {name: getattr(st, name) for name, _ in RequestStats._fields_}
        return totals[:n], (counts[:n] if counts is not None else None), agg, stats

    def poll(self, ticket: int):
        """None while the request is queued or running, else what wait() returns."""
        st = RequestStats()
        rc = lib().zg_server_poll(self._h, ticket, C.byref(st))
        if rc == 0:
            return None
        buf, off, n, totals, counts, agg = self._live.pop(ticket)
        check(0 if rc == 1 else rc)
        stats = {name: getattr(st, name) for name, _ in RequestStats._fields_}
        return totals[:n], (counts[:n] if counts is not None else None), agg, stats

    def stats(self) -> dict:
        s = ServerStats()
        check(lib().zg_server_get_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            lib().zg_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Builder:
    """In-memory index construction (power_merge::build, storage.rs:500-882)."""

    def __init__(self, **config):
        self._cfg = build_config(**config)
        h = C.c_void_p()
        check(lib().zg_builder_create(C.byref(self._cfg), C.byref(h)))
        self._h = h

    def add_color(self, name: str, records) -> None:
        """Add one dataset from a list of sequences (bytes) or (buffer, offsets)."""
        if isinstance(records, tuple):
            buf, off = _as_reads(*records)
        else:
            buf, off = pack_reads(records)
        check(lib().zg_builder_add_color(self._h, name.encode(), buf.ctypes.data_as(C.c_void_p),
                                         _u64p(off), len(off) - 1))

    def finish(self) -> Index:
        h = C.c_void_p()
        rc = lib().zg_builder_finish(self._h, C.byref(h))
        self._h = None
        check(rc)
        return Index(h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().zg_builder_destroy(self._h)
            self._h = None


def build(fof: str, output: str, err_fd: int = 2, **config) -> None:
    """zorindex build --fof FOF --output OUT (naive planner, width column order)."""
    cfg = build_config(**config)
    check(lib().zg_build_command(C.byref(cfg), str(fof).encode(), str(output).encode(), err_fd))


def fuse_many_indexes(inputs, output: str, err_fd: int = 2) -> None:
    """fusion.rs:233-480: concatenate compatible indexes into `output`."""
    arr = (C.c_char_p * len(inputs))(*[str(p).encode() for p in inputs])
    check(lib().zg_fusion_command(arr, len(inputs), str(output).encode(), err_fd))


def fuse_indexes(first: str, second: str, output: str, err_fd: int = 2) -> None:
    fuse_many_indexes([first, second], output, err_fd)


def query_command_with_memory(directory: str, query: str, min_ratio: float = 0.0,
                              recover: bool = True, memory: str = "subindex",
                              out_fd: int = 1, err_fd: int = 2, hbm_budget: int = 0) -> None:
    """storage.rs:3731-3767: aggregate query, TSV to out_fd, diagnostics to err_fd."""
    if memory not in QUERY_MEMORY:
        raise ValueError(f"invalid query memory mode {memory}")
    check(lib().zg_query_command_ex(str(directory).encode(), str(query).encode(),
                                    float(min_ratio), int(bool(recover)),
                                    QUERY_MEMORY.index(memory), int(hbm_budget), out_fd, err_fd))


def query_command_with_recovery(directory: str, query: str, min_ratio: float,
                                recover: bool, **kw) -> None:
    query_command_with_memory(directory, query, min_ratio, recover, **kw)


def query_command(directory: str, query: str, min_ratio: float = 0.0, **kw) -> None:
    query_command_with_memory(directory, query, min_ratio, True, **kw)


def query_records_command_with_memory(directory: str, query: str, min_ratio: float = 0.0,
                                      recover: bool = True, batch_size: int = 4096,
                                      memory: str = "subindex", format: str = "tsv",
                                      out_fd: int = 1, err_fd: int = 2,
                                      hbm_budget: int = 0) -> None:
    """storage.rs:3626-3720: per-record output (tsv or compact)."""
    if memory not in QUERY_MEMORY:
        raise ValueError(f"invalid query memory mode {memory}")
    check(lib().zg_query_records_command_ex(str(directory).encode(), str(query).encode(),
                                            float(min_ratio), int(bool(recover)),
                                            int(batch_size), RECORD_FORMATS[format],
                                            QUERY_MEMORY.index(memory), int(hbm_budget),
                                            out_fd, err_fd))


def query_records_command(directory: str, query: str, min_ratio: float, recover: bool,
                          batch_size: int, format: str = "tsv", **kw) -> None:
    query_records_command_with_memory(directory, query, min_ratio, recover, batch_size,
                                      format=format, **kw)


def info_command(directory: str, out_fd: int = 1) -> None:
    check(lib().zg_info_command(str(directory).encode(), out_fd))
