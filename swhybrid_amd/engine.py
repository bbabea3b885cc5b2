"""Python front end of the HIP search library.

    eng = Engine("blosum62", gap_open=10, gap_extend=2)
    db = eng.load_db(seqdb)                 # packed and resident in HBM
    scores, stats = eng.search(db, query)   # exact int32 scores, caller order
    top = topk(scores, 20)                  # (score desc, index asc)

All work happens in libswh.so; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _native as N
from .alphabet import matrix as named_matrix
from .datagen import SeqDB


def _as_matrix(m) -> np.ndarray:
    if isinstance(m, str):
        m = named_matrix(m)
    m = np.ascontiguousarray(np.asarray(m, dtype=np.int32))
    if m.shape != (32, 32):
        raise ValueError("matrix must be 32x32")
    return m


def _as_query(q) -> np.ndarray:
    q = np.ascontiguousarray(np.asarray(q, dtype=np.uint8))
    if q.ndim != 1:
        raise ValueError("query must be a 1-D code array")
    return q


_PROFILE_MODES = {0: 0, "auto": 0, 1: 1, "single": 1, 2: 2, "pair": 2}


def _opts(strip_width=0, overflow_threshold=0, group_lanes=0, group_width=0, use_graph=0,
          profile_mode=0):
    if profile_mode not in _PROFILE_MODES:
        raise ValueError(f"profile_mode {profile_mode!r}: use 'auto', 'single' or 'pair'")
    return N.SwhSearchOpts(int(strip_width), int(overflow_threshold), int(group_lanes),
                           int(group_width), int(use_graph), _PROFILE_MODES[profile_mode])


class Database:
    def __init__(self, engine: "Engine", handle, n: int, residues: int):
        self._engine = engine
        self.handle = handle
        self.n = n
        self.residues = residues

    def info(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.check(N.lib().swh_db_info(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return {"nseq": a.value, "residues": b.value, "device_bytes": c.value}

    def close(self):
        if self.handle:
            N.check(N.lib().swh_db_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    def __init__(self, matrix="blosum62", gap_open=10, gap_extend=2, device=0):
        self.matrix = _as_matrix(matrix)
        self.gap_open = int(gap_open)
        self.gap_extend = int(gap_extend)
        h = C.c_void_p()
        N.check(N.lib().swh_ctx_create(int(device), N.ptr(self.matrix), self.gap_open,
                                       self.gap_extend, C.byref(h)))
        self.handle = h
        self._dbs = weakref.WeakSet()   # open databases, closed before the context

    def set_scoring(self, matrix, gap_open, gap_extend):
        m = _as_matrix(matrix)
        N.check(N.lib().swh_ctx_set_scoring(self.handle, N.ptr(m), int(gap_open),
                                            int(gap_extend)))
        self.matrix, self.gap_open, self.gap_extend = m, int(gap_open), int(gap_extend)

    def load_db(self, db: SeqDB, group_min_len: int = 0, queries_per_launch: int = 0) -> Database:
        """Pack and upload; sequences of at least group_min_len residues are
        stored as pairs for group mode (0 = default: total residues / (2560 x
        CUs), at least 2048; -1 = never).  With the default, a database of
        fewer than 1024 sequences per CU also gets a layout with pairs from
        128 residues for launches of few queries, unless queries_per_launch
        says which one the caller's launches need."""
        res = np.ascontiguousarray(db.residues, dtype=np.uint8)
        off = np.ascontiguousarray(db.offsets, dtype=np.uint64)
        lens = np.ascontiguousarray(db.lengths, dtype=np.uint32)
        h = C.c_void_p()
        o = N.SwhDbOpts(0xFFFFFFFF if group_min_len < 0 else int(group_min_len),
                        int(queries_per_launch))
        N.check(N.lib().swh_db_create(self.handle, N.ptr(res), N.ptr(off), N.ptr(lens),
                                      db.n, C.byref(o), C.byref(h)))
        d = Database(self, h, db.n, db.total)
        self._dbs.add(d)
        return d

    def pinned_scores(self, n, nq=1):
        """A page-locked int32 array of nq x n (1-d for nq = 1) for the out=
        argument of search / search_batch: the device copies scores straight
        into it.  Freed when the array is."""
        return N.pinned_array(n * nq, np.int32).reshape((nq, n) if nq > 1 else (n,))

    def search(self, db: Database, query, out=None, **opts):
        q = _as_query(query)
        if out is None:
            out = np.empty(db.n, dtype=np.int32)
        elif out.dtype != np.int32 or out.size < db.n or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous int32 array of at least db.n entries")
        st = N.SwhStats()
        o = _opts(**opts)
        N.check(N.lib().swh_search(self.handle, db.handle, N.ptr(q), len(q), C.byref(o),
                                   N.ptr(out), C.byref(st)))
        return out, st.as_dict()

    def search_batch(self, db: Database, queries, out=None, **opts):
        qs = [_as_query(q) for q in queries]
        arr = (C.c_void_p * len(qs))(*[q.ctypes.data for q in qs])
        lens = np.array([len(q) for q in qs], dtype=np.int32)
        if out is None:
            out = np.empty((len(qs), db.n), dtype=np.int32)
        elif (out.dtype != np.int32 or out.size < len(qs) * db.n
              or not out.flags.c_contiguous):
            raise ValueError("out must be a contiguous int32 array of at least nq x db.n")
        st = N.SwhStats()
        o = _opts(**opts)
        N.check(N.lib().swh_search_batch(self.handle, db.handle, arr, N.ptr(lens), len(qs),
                                         C.byref(o), N.ptr(out), C.byref(st)))
        return out, st.as_dict()

    def autotune(self, db: SeqDB, bucket_qlens, sample_every=8, queries_per_bucket=4, reps=1):
        """Measure launch shapes per query-length bucket on every
        sample_every-th sequence of db in length order, install the table on
        this engine and return it with the best group_min_len for load_db."""
        res = np.ascontiguousarray(db.residues, dtype=np.uint8)
        off = np.ascontiguousarray(db.offsets, dtype=np.uint64)
        lens = np.ascontiguousarray(db.lengths, dtype=np.uint32)
        b = np.ascontiguousarray(np.asarray(bucket_qlens, dtype=np.int32))
        out = N.SwhTuneResult()
        N.check(N.lib().swh_autotune(self.handle, N.ptr(res), N.ptr(off), N.ptr(lens), db.n,
                                     int(sample_every), N.ptr(b), len(b),
                                     int(queries_per_bucket), int(reps), C.byref(out)))
        return {
            "group_min_len": int(out.group_min_len),
            "entries": [out.entries[i].as_dict() for i in range(out.n_entries)],
            "seconds": float(out.seconds),
            "sample_seqs": int(out.sample_seqs),
            "sample_residues": int(out.sample_residues),
        }

    def set_tuning(self, entries):
        """Install a tuning table (list of dicts as autotune returns; [] clears)."""
        arr = (N.SwhTuneEntry * max(1, len(entries)))()
        for i, e in enumerate(entries):
            t = arr[i]
            for f in ("qlen_max", "strip_width", "profile_mode", "group_lanes", "group_width"):
                setattr(t, f, int(e[f]))
            cost = list(e["lane_cost"])
            if len(cost) != N.SWH_TUNE_LANE_CONFIGS:
                raise ValueError("lane_cost needs one value per lane configuration")
            for k, v in enumerate(cost):
                t.lane_cost[k] = float(v)
        N.check(N.lib().swh_set_tuning(self.handle, arr, len(entries)))

    def get_tuning(self):
        arr = (N.SwhTuneEntry * N.SWH_TUNE_MAX)()
        n = C.c_int(0)
        N.check(N.lib().swh_get_tuning(self.handle, arr, N.SWH_TUNE_MAX, C.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def close(self):
        """Closes the databases this engine loaded, then the context (the C
        ABI refuses to destroy a context whose databases still exist)."""
        if self.handle:
            for d in list(self._dbs):
                d.close()
            N.check(N.lib().swh_ctx_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching search service on a resident database (a loop
    thread in libswh.so).  submit() returns a request id at once; wait()
    blocks for its top-k (and optionally its full score vector)."""

    def __init__(self, engine: Engine, db: Database, max_inflight=0, memory_budget=0,
                 slices=0, max_qlen=0, max_k=0, round_ms=0.0, wait_weight=0.0, queue="auto",
                 **search_opts):
        """queue: "live" (one persistent launch fed while it runs, the
        default on a database of one layout), "rounds" (one launch per round;
        slices / round_ms shape them) or "auto"."""
        o = N.SwhServerOpts(int(max_inflight), int(memory_budget), int(slices), int(max_qlen),
                            int(max_k), _opts(**search_opts), float(round_ms), float(wait_weight),
                            N.QUEUE[queue])
        h = C.c_void_p()
        N.check(N.lib().swh_server_create(engine.handle, db.handle, C.byref(o), C.byref(h)))
        self.handle = h
        self.db = db
        self._k = {}
        self._want = {}

    def submit(self, query, k=10, want_scores=False) -> int:
        q = _as_query(query)
        rid = C.c_uint64()
        N.check(N.lib().swh_server_submit(self.handle, N.ptr(q), len(q), int(k),
                                          int(bool(want_scores)), C.byref(rid)))
        self._k[rid.value] = int(k)
        self._want[rid.value] = bool(want_scores)
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1):
        k = self._k.pop(rid)
        want = self._want.pop(rid)
        ts = np.zeros(max(k, 1), dtype=np.int32)
        ti = np.zeros(max(k, 1), dtype=np.int64)
        sc = np.zeros(self.db.n, dtype=np.int32) if want else None
        info = N.SwhRequestInfo()
        rc = N.lib().swh_server_wait(self.handle, C.c_uint64(rid), int(timeout_ms), N.ptr(ts),
                                     N.ptr(ti), None if sc is None else N.ptr(sc), C.byref(info))
        if rc != 0:
            self._k[rid] = k
            self._want[rid] = want
            N.check(rc)
        return {"top_score": ts[:k], "top_index": ti[:k], "scores": sc, "info": info.as_dict()}

    def stats(self):
        st = N.SwhServerStats()
        N.check(N.lib().swh_server_get_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.handle:
            N.check(N.lib().swh_server_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def topk(scores, k: int):
    """Top-k of a score vector, (score desc, index asc), padded with (0, 0)."""
    s = np.ascontiguousarray(scores, dtype=np.int32)
    os_ = np.zeros(k, dtype=np.int32)
    oi = np.zeros(k, dtype=np.int64)
    N.check(N.lib().swh_topk(N.ptr(s), s.shape[0], int(k), N.ptr(os_), N.ptr(oi)))
    return os_, oi


def device_count() -> int:
    n = C.c_int(0)
    N.check(N.lib().swh_device_count(C.byref(n)))
    return n.value


def pool_search(db: SeqDB, query, matrix="blosum62", gap_open=10, gap_extend=2, k=10,
                chunk_bytes=0, n_workers=1, all_scores=False):
    """Run one query through the C++ worker pipeline (DataPool -> HIPWorker
    threads -> RecalcWorker -> merged top-k), as the reference's main() does.
    Returns dict(top_score, top_index, top_pool_index, scores, seconds, gcups)."""
    m = _as_matrix(matrix)
    q = _as_query(query)
    res = np.ascontiguousarray(db.residues, dtype=np.uint8)
    off = np.ascontiguousarray(db.offsets, dtype=np.uint64)
    lens = np.ascontiguousarray(db.lengths, dtype=np.uint32)
    ts = np.zeros(k, dtype=np.int32)
    ti = np.zeros(k, dtype=np.int64)
    tp = np.zeros(k, dtype=np.int64)
    sc = np.zeros(db.n, dtype=np.int32) if all_scores else None
    secs, gc = C.c_double(0), C.c_double(0)
    L = N.host_lib()
    rc = L.swhy_pool_search(N.ptr(res), N.ptr(off), N.ptr(lens), db.n, N.ptr(q), len(q), N.ptr(m),
                            int(gap_open), int(gap_extend), int(k), int(chunk_bytes),
                            int(n_workers), N.ptr(ts), N.ptr(ti), N.ptr(tp),
                            None if sc is None else N.ptr(sc), C.byref(secs), C.byref(gc))
    if rc != 0:
        msg = L.swhy_last_error()
        raise N.SwhError(f"pool search failed ({rc}): {msg.decode() if msg else ''}")
    return {"top_score": ts, "top_index": ti, "top_pool_index": tp, "scores": sc,
            "seconds": secs.value, "gcups": gc.value}


def cpu_search(db: SeqDB, query, matrix="blosum62", gap_open=10, gap_extend=2, threads=0,
               query_segment=0):
    """The host's AVX2 scan alone (int8 pass, int16 for lanes that may have
    saturated) over the whole database with `threads` threads (0 = every
    core).  Returns (scores, scan seconds, pack seconds); a score beyond
    int16 is -1."""
    m = _as_matrix(matrix)
    q = _as_query(query)
    res = np.ascontiguousarray(db.residues, dtype=np.uint8)
    off = np.ascontiguousarray(db.offsets, dtype=np.uint64)
    lens = np.ascontiguousarray(db.lengths, dtype=np.uint32)
    out = np.zeros(db.n, dtype=np.int32)
    secs, pack = C.c_double(0), C.c_double(0)
    L = N.host_lib()
    rc = L.swhy_cpu_search(N.ptr(res), N.ptr(off), N.ptr(lens), db.n, N.ptr(q), len(q), N.ptr(m),
                           int(gap_open), int(gap_extend), int(threads), int(query_segment),
                           N.ptr(out), C.byref(secs), C.byref(pack))
    if rc != 0:
        msg = L.swhy_last_error()
        raise N.SwhError(f"cpu search failed ({rc}): {msg.decode() if msg else ''}")
    return out, secs.value, pack.value


def hybrid_search(db: SeqDB, query, matrix="blosum62", gap_open=10, gap_extend=2, k=10,
                  hip_workers=1, cpu_threads=0, chunk_bytes=0, cpu_chunk_bytes=0,
                  all_scores=False, first_chunk_bytes=0):
    """The worker pipeline with HIPWorkers and a CPU worker (AVX2, cpu_threads
    threads) pulling chunks from one DataPool.  chunk_bytes / first_chunk_bytes:
    a HIPWorker's chunk schedule (its first chunk, each later one twice the one
    before, up to chunk_bytes; both 0 = the default schedule by query length).
    Returns dict(top_score, top_index, scores, report); report names the
    schedule that ran."""
    m = _as_matrix(matrix)
    q = _as_query(query)
    res = np.ascontiguousarray(db.residues, dtype=np.uint8)
    off = np.ascontiguousarray(db.offsets, dtype=np.uint64)
    lens = np.ascontiguousarray(db.lengths, dtype=np.uint32)
    ts = np.zeros(k, dtype=np.int32)
    ti = np.zeros(k, dtype=np.int64)
    sc = np.zeros(db.n, dtype=np.int32) if all_scores else None
    o = N.SwhyPoolOpts(int(k), int(hip_workers), int(cpu_threads), int(chunk_bytes),
                       int(cpu_chunk_bytes), int(first_chunk_bytes))
    rep = N.SwhyPoolReport()
    L = N.host_lib()
    rc = L.swhy_pool_search_ex(N.ptr(res), N.ptr(off), N.ptr(lens), db.n, N.ptr(q), len(q),
                               N.ptr(m), int(gap_open), int(gap_extend), C.byref(o), N.ptr(ts),
                               N.ptr(ti), None if sc is None else N.ptr(sc), C.byref(rep))
    if rc != 0:
        msg = L.swhy_last_error()
        raise N.SwhError(f"hybrid search failed ({rc}): {msg.decode() if msg else ''}")
    return {"top_score": ts, "top_index": ti, "scores": sc, "report": rep.as_dict()}
