"""ctypes binding of libswh.so (the C ABI declared in include/swh.h).

The shared object is built in-tree (swhybrid_amd/libswh.so).  Loading it
never falls back to anything: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswh.so")
HOST_LIB_PATH = os.path.join(HERE, "libswhybrid_host.so")

SWH_OK = 0

# Every symbol include/swh.h declares (tests check the library exports them).
EXPORTS = (
    "swh_last_error", "swh_version", "swh_device_count", "swh_ctx_create",
    "swh_ctx_destroy", "swh_ctx_set_scoring", "swh_db_create", "swh_db_destroy",
    "swh_db_info", "swh_search", "swh_search_batch", "swh_topk",
    "swh_server_create", "swh_server_submit", "swh_server_wait", "swh_server_get_stats",
    "swh_server_destroy", "swh_autotune", "swh_set_tuning", "swh_get_tuning",
    "swh_pinned_alloc", "swh_pinned_free", "swh_search_noalloc",
)

SWH_PROFILE_AUTO, SWH_PROFILE_SINGLE, SWH_PROFILE_PAIR = 0, 1, 2
SWH_TUNE_MAX = 16
SWH_TUNE_LANE_CONFIGS = 5
# (strip width, pair table) of the lane_cost columns, as in include/swh.h
LANE_CONFIGS = ((16, True), (24, True), (16, False), (24, False), (32, False))


class SwhStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("rescore_ms", C.c_double),
        ("total_ms", C.c_double),
        ("cells", C.c_uint64),
        ("cells_group", C.c_uint64),
        ("padded_cells", C.c_uint64),
        ("n_overflow", C.c_uint64),
        ("n_launches", C.c_uint64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SwhSearchOpts(C.Structure):
    _fields_ = [
        ("strip_width", C.c_int),
        ("overflow_threshold", C.c_int),
        ("group_lanes", C.c_int),
        ("group_width", C.c_int),
        ("use_graph", C.c_int),
        ("profile_mode", C.c_int),
    ]


class SwhServerOpts(C.Structure):
    _fields_ = [
        ("max_inflight", C.c_int),
        ("memory_budget", C.c_uint64),
        ("slices", C.c_int),
        ("max_qlen", C.c_int),
        ("max_k", C.c_int),
        ("search", SwhSearchOpts),
        ("round_ms", C.c_double),
        ("wait_weight", C.c_double),
        ("queue", C.c_int),
    ]


QUEUE = {"auto": 0, "live": 1, "rounds": 2}


class SwhRequestInfo(C.Structure):
    _fields_ = [
        ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double),
        ("latency_ms", C.c_double), ("cells", C.c_uint64), ("rounds", C.c_int),
        ("status", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SwhServerStats(C.Structure):
    _fields_ = [
        ("rounds", C.c_uint64), ("completed", C.c_uint64), ("cells", C.c_uint64),
        ("busy_ms", C.c_double), ("peak_inflight", C.c_int), ("waiting", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SwhDbOpts(C.Structure):
    _fields_ = [("group_min_len", C.c_uint32), ("queries_per_launch", C.c_uint32)]


class SwhTuneEntry(C.Structure):
    _fields_ = [
        ("qlen_max", C.c_int), ("strip_width", C.c_int), ("profile_mode", C.c_int),
        ("group_lanes", C.c_int), ("group_width", C.c_int),
        ("lane_cost", C.c_float * SWH_TUNE_LANE_CONFIGS),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "lane_cost"}
        d["lane_cost"] = [float(v) for v in self.lane_cost]
        return d


class SwhTuneResult(C.Structure):
    _fields_ = [
        ("group_min_len", C.c_uint32), ("n_entries", C.c_int),
        ("entries", SwhTuneEntry * SWH_TUNE_MAX), ("seconds", C.c_double),
        ("sample_seqs", C.c_uint64), ("sample_residues", C.c_uint64),
    ]


class SwhError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libswh.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SwhError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "or `make -C swhybrid_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.swh_last_error.restype = C.c_char_p
    L.swh_last_error.argtypes = []
    L.swh_version.restype = C.c_int
    L.swh_device_count.argtypes = [C.POINTER(C.c_int)]
    L.swh_ctx_create.argtypes = [C.c_int, vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.swh_ctx_destroy.argtypes = [vp]
    L.swh_ctx_set_scoring.argtypes = [vp, vp, C.c_int, C.c_int]
    L.swh_db_create.argtypes = [vp, vp, vp, vp, C.c_uint64, C.POINTER(SwhDbOpts), C.POINTER(vp)]
    L.swh_db_destroy.argtypes = [vp]
    L.swh_db_info.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                              C.POINTER(C.c_uint64)]
    L.swh_search.argtypes = [vp, vp, vp, C.c_int, C.POINTER(SwhSearchOpts), vp,
                             C.POINTER(SwhStats)]
    L.swh_search_noalloc.argtypes = [vp, vp, vp, C.c_int, C.POINTER(SwhSearchOpts), vp,
                                     C.POINTER(SwhStats)]
    L.swh_search_batch.argtypes = [vp, vp, vp, vp, C.c_int, C.POINTER(SwhSearchOpts), vp,
                                   C.POINTER(SwhStats)]
    L.swh_topk.argtypes = [vp, C.c_uint64, C.c_int, vp, vp]
    L.swh_server_create.argtypes = [vp, vp, C.POINTER(SwhServerOpts), C.POINTER(vp)]
    L.swh_server_submit.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint64)]
    L.swh_server_wait.argtypes = [vp, C.c_uint64, C.c_int, vp, vp, vp, C.POINTER(SwhRequestInfo)]
    L.swh_server_get_stats.argtypes = [vp, C.POINTER(SwhServerStats)]
    L.swh_server_destroy.argtypes = [vp]
    L.swh_autotune.argtypes = [vp, vp, vp, vp, C.c_uint64, C.c_int, vp, C.c_int, C.c_int,
                               C.c_int, C.POINTER(SwhTuneResult)]
    L.swh_set_tuning.argtypes = [vp, C.POINTER(SwhTuneEntry), C.c_int]
    L.swh_pinned_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.swh_pinned_free.argtypes = [vp]
    L.swh_get_tuning.argtypes = [vp, C.POINTER(SwhTuneEntry), C.c_int, C.POINTER(C.c_int)]
    for name in EXPORTS:
        if name != "swh_last_error" and name != "swh_version":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


HOST_EXPORTS = ("swhy_last_error", "swhy_pool_search", "swhy_dbfile_make", "swhy_dbfile_open",
                "swhy_dbfile_info", "swhy_dbfile_arrays", "swhy_dbfile_title",
                "swhy_dbfile_close", "swhy_dbfile_pool_search", "swhy_fasta_open",
                "swhy_fasta_count", "swhy_fasta_record", "swhy_fasta_close",
                "swhy_pool_search_ex", "swhy_dbfile_pool_search_ex", "swhy_cpu_search")


class SwhyPoolOpts(C.Structure):
    _fields_ = [("topk", C.c_int), ("hip_workers", C.c_int), ("cpu_threads", C.c_int),
                ("chunk_bytes", C.c_uint64), ("cpu_chunk_bytes", C.c_uint64),
                ("first_chunk_bytes", C.c_uint64)]


class SwhyPoolReport(C.Structure):
    _fields_ = [("seconds", C.c_double), ("gcups", C.c_double), ("hip_residues", C.c_uint64),
                ("cpu_residues", C.c_uint64), ("hip_busy_s", C.c_double),
                ("cpu_busy_s", C.c_double), ("recalc", C.c_uint64),
                ("wall_seconds", C.c_double), ("dev_first_fill_s", C.c_double),
                ("dev_aux_fill_s", C.c_double), ("dev_pack_s", C.c_double),
                ("dev_search_s", C.c_double), ("dev_aux_wait_s", C.c_double),
                ("dev_kernel_s", C.c_double), ("dev_chunks", C.c_int),
                ("chunk_bytes", C.c_uint64), ("first_chunk_bytes", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}
_host = None


def host_lib():
    """Load libswhybrid_host.so, the C++ worker pipeline (DataPool, HIPWorker,
    RecalcWorker, WorkerPool) built on top of libswh.so."""
    global _host
    if _host is not None:
        return _host
    lib()
    if not os.path.exists(HOST_LIB_PATH):
        raise SwhError(f"{HOST_LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(HOST_LIB_PATH)
    vp = C.c_void_p
    L.swhy_last_error.restype = C.c_char_p
    L.swhy_pool_search.restype = C.c_int
    L.swhy_pool_search.argtypes = [vp, vp, vp, C.c_uint64, vp, C.c_int, vp, C.c_int, C.c_int,
                                   C.c_int, C.c_uint64, C.c_int, vp, vp, vp, vp,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
    u64p = C.POINTER(C.c_uint64)
    L.swhy_dbfile_make.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64, u64p, u64p]
    L.swhy_dbfile_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.swhy_dbfile_info.argtypes = [vp, u64p, u64p, u64p, u64p]
    L.swhy_dbfile_arrays.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    L.swhy_dbfile_title.argtypes = [vp, C.c_uint64, C.c_char_p, C.c_uint64, u64p]
    L.swhy_dbfile_close.argtypes = [vp]
    L.swhy_dbfile_pool_search.argtypes = [vp, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                          C.c_uint64, C.c_int, vp, vp, vp,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.swhy_fasta_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.swhy_fasta_count.argtypes = [vp, u64p]
    L.swhy_fasta_record.argtypes = [vp, C.c_uint64, C.POINTER(vp), u64p,
                                    C.POINTER(C.c_char_p)]
    L.swhy_fasta_close.argtypes = [vp]
    L.swhy_pool_search_ex.argtypes = [vp, vp, vp, C.c_uint64, vp, C.c_int, vp, C.c_int, C.c_int,
                                      C.POINTER(SwhyPoolOpts), vp, vp, vp,
                                      C.POINTER(SwhyPoolReport)]
    L.swhy_dbfile_pool_search_ex.argtypes = [vp, vp, C.c_int, vp, C.c_int, C.c_int,
                                             C.POINTER(SwhyPoolOpts), vp, vp, vp,
                                             C.POINTER(SwhyPoolReport)]
    L.swhy_cpu_search.argtypes = [vp, vp, vp, C.c_uint64, vp, C.c_int, vp, C.c_int, C.c_int,
                                  C.c_int, C.c_int, vp, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    for name in HOST_EXPORTS[1:]:
        getattr(L, name).restype = C.c_int
    _host = L
    return L


def check(rc: int):
    if rc != SWH_OK:
        msg = lib().swh_last_error()
        raise SwhError(f"libswh error {rc}: {msg.decode() if msg else ''}")


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class _Pinned:
    """Page-locked host memory from swh_pinned_alloc, exposed to numpy through
    the array interface and freed with the last array that uses it."""

    def __init__(self, n, dtype):
        dt = np.dtype(dtype)
        self.addr = C.c_void_p()
        check(lib().swh_pinned_alloc(max(1, n) * dt.itemsize, C.byref(self.addr)))
        self.__array_interface__ = {"shape": (n,), "typestr": dt.str,
                                    "data": (self.addr.value, False), "version": 3}

    def __del__(self):
        if getattr(self, "addr", None) is not None and self.addr.value:
            lib().swh_pinned_free(self.addr)
            self.addr = None


def pinned_array(n, dtype=np.int32) -> np.ndarray:
    """A 1-d array of n elements in page-locked host memory."""
    return np.asarray(_Pinned(n, dtype))
