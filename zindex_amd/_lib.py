"""ctypes binding of libzindex_amd.so (declarations mirror include/zindex_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the query path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libzindex_amd.so"

ZG_OK = 0
ZG_ERR_INVALID = -1
ZG_ERR_IO = -2
ZG_ERR_FORMAT = -3
ZG_ERR_DEVICE = -4
ZG_ERR_NOMEM = -5
ZG_ERR_UNSUPPORTED = -6
ZG_ERR_BUDGET = -7

ZG_QUERY_RECORDS = 0x1
ZG_QUERY_NO_RECOVERY = 0x2
ZG_MEMORY_ALL = 0
ZG_MEMORY_SUBINDEX = 1

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class IndexMeta(C.Structure):
    _fields_ = [
        ("version", C.c_uint32), ("k", C.c_uint32), ("hashes", C.c_uint32),
        ("fingerprint_bits", C.c_uint32), ("index_mode", C.c_uint32),
        ("minimizer_size", C.c_uint32), ("minimizer_scheme", C.c_uint32),
        ("n_colors", C.c_uint32), ("n_banks", C.c_uint32), ("n_classes", C.c_uint32),
        ("seed", C.c_uint64), ("segment_length", C.c_uint64),
        ("abandoned_recovery", C.c_int32), ("reserved", C.c_uint32),
        ("abandoned_hashes", C.c_uint64), ("abandoned_origins", C.c_uint64),
        ("payload_bytes", C.c_uint64),
    ]


class BuildConfig(C.Structure):
    _fields_ = [
        ("k", C.c_uint32), ("hashes", C.c_uint32), ("fingerprint_bits", C.c_uint32),
        ("index_mode", C.c_uint32), ("minimizer_size", C.c_uint32),
        ("segment_length", C.c_uint32), ("seed", C.c_uint64), ("slot_scale", C.c_double),
        ("max_subindex_size", C.c_uint64), ("merge_group_size", C.c_uint32),
        ("threads", C.c_uint32), ("drop_abandoned", C.c_int32),
        ("modimizer_sampling", C.c_uint32), ("device", C.c_uint32),
        ("homopolymer_compression", C.c_uint32), ("findere_z", C.c_uint32),
        ("seed_pattern", C.c_char * 32), ("randstrobemers", C.c_uint32),
        ("strobe_min", C.c_uint32), ("strobe_max", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [("span", C.c_uint32), ("tiles_per_wave", C.c_uint32),
                ("prefetch", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("nontemporal", C.c_uint32), ("sweep", C.c_uint32), ("graph", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [
        ("n_reads", C.c_uint64), ("bases", C.c_uint64), ("seeds", C.c_uint64),
        ("lookups", C.c_uint64), ("algorithmic_bytes", C.c_uint64),
        ("last_lookup_ms", C.c_double), ("last_seed_ms", C.c_double),
        ("last_total_ms", C.c_double), ("tuning", Tuning), ("sweep_folded", C.c_uint64),
        ("sweep_launches", C.c_uint64), ("sweep_lookup_ms", C.c_double),
        ("sweep_sort_ms", C.c_double), ("sweep_sum_ms", C.c_double),
        ("sweep_stage_launches", C.c_uint64), ("sweep_algorithmic_bytes", C.c_uint64),
        ("sweep_hit_row_bytes", C.c_uint64), ("sweep_plane_banks", C.c_uint64),
        ("sweep_recovery_ms", C.c_double),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("slots", C.c_uint32), ("max_requests", C.c_uint32), ("slot_reads", C.c_uint64),
        ("slot_bases", C.c_uint64), ("hbm_budget", C.c_uint64), ("flags", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class RequestStats(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64), ("batch_id", C.c_uint64), ("batch_requests", C.c_uint32),
        ("status", C.c_int32), ("n_reads", C.c_uint64), ("seeds", C.c_uint64),
        ("queued_ms", C.c_double), ("service_ms", C.c_double), ("latency_ms", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("requests", C.c_uint64), ("batches", C.c_uint64), ("reads", C.c_uint64),
        ("seeds", C.c_uint64), ("max_batch_requests", C.c_uint32), ("slots", C.c_uint32),
        ("lookup_ms", C.c_double), ("device_bytes", C.c_uint64),
        ("sweep_batches", C.c_uint64),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "zg_last_error": (C.c_char_p, []),
    "zg_version": (C.c_char_p, []),
    "zg_index_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "zg_index_close": (C.c_int, [C.c_void_p]),
    "zg_index_get_meta": (C.c_int, [C.c_void_p, C.POINTER(IndexMeta)]),
    "zg_index_color": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_char_p), u64p]),
    "zg_index_load": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64]),
    "zg_index_load_ex": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_uint32]),
    "zg_index_bank_payload": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(u8p), u64p]),
    "zg_index_manifest": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), u64p]),
    "zg_index_recovery": (C.c_int, [C.c_void_p, C.POINTER(u32p), C.POINTER(u64p),
                                    C.POINTER(u32p), u64p]),
    "zg_query_reads": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.c_uint64, C.c_uint32,
                                 u64p, u32p, u64p]),
    "zg_query_reads_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.c_uint64, C.c_uint32,
                                         u64p, u32p, u64p]),
    "zg_extract_seeds": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.c_uint64, u64p, u64p,
                                   u64p, C.c_uint64]),
    "zg_query_command": (C.c_int, [C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_int,
                                   C.c_int]),
    "zg_query_records_command": (C.c_int, [C.c_char_p, C.c_char_p, C.c_double, C.c_int,
                                           C.c_uint64, C.c_int, C.c_int, C.c_int]),
    "zg_query_command_ex": (C.c_int, [C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_uint32,
                                      C.c_uint64, C.c_int, C.c_int]),
    "zg_query_records_command_ex": (C.c_int, [C.c_char_p, C.c_char_p, C.c_double, C.c_int,
                                              C.c_uint64, C.c_int, C.c_uint32, C.c_uint64,
                                              C.c_int, C.c_int]),
    "zg_info_command": (C.c_int, [C.c_char_p, C.c_int]),
    "zg_build_config_default": (None, [C.POINTER(BuildConfig)]),
    "zg_fusion_command": (C.c_int, [C.POINTER(C.c_char_p), C.c_uint32, C.c_char_p, C.c_int]),
    "zg_build_command": (C.c_int, [C.POINTER(BuildConfig), C.c_char_p, C.c_char_p, C.c_int]),
    "zg_builder_create": (C.c_int, [C.POINTER(BuildConfig), C.POINTER(C.c_void_p)]),
    "zg_builder_add_color": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, u64p, C.c_uint64]),
    "zg_builder_finish": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "zg_index_write": (C.c_int, [C.c_void_p, C.c_char_p]),
    "zg_builder_destroy": (C.c_int, [C.c_void_p]),
    "zg_batch_create": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.c_uint64, C.c_uint32,
                                  C.POINTER(C.c_void_p)]),
    "zg_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zg_batch_run_plain": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64]),
    "zg_batch_reseed": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zg_batch_results": (C.c_int, [C.c_void_p, u64p, u32p, u64p]),
    "zg_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "zg_batch_destroy": (C.c_int, [C.c_void_p]),
    "zg_index_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "zg_index_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "zg_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Tuning)]),
    "zg_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "zg_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "zg_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.c_uint64, u64p, u32p, u64p,
                                   u64p]),
    "zg_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestStats)]),
    "zg_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestStats)]),
    "zg_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "zg_server_destroy": (C.c_int, [C.c_void_p]),
    "zg_batch_store_probe": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "zg_gather_probe": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64,
                                  C.POINTER(C.c_double)]),
    "zg_gather_probe_keys": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64,
                                       C.POINTER(C.c_double)]),
    "zg_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "zg_host_free": (C.c_int, [C.c_void_p]),
}


class ZindexError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -m zindex_amd.buildlib)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != ZG_OK:
        msg = lib().zg_last_error().decode(errors="replace")
        raise ZindexError(rc, msg)
