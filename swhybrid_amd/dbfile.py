"""The reference's on-disk database (<base>.seq/.map/.title) and FASTA input,
through libswhybrid_host.so.

    make("db.fasta", "out/db")            # the DBmaker
    with DbFile("out/db") as f:
        seqdb = f.seqdb()                 # SeqDB over the mapped .seq (no copy)
        f.title(i)                        # '>' line of sequence i
    records = read_fasta("queries.fa")    # [(title, codes)], every record

A sequence's index is its position in .seq (the fillers that complete the
final batch included), the index the reference's results and titles use.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .datagen import SeqDB


def _check(rc):
    if rc != 0:
        msg = N.host_lib().swhy_last_error()
        raise N.SwhError(msg.decode() if msg else f"error {rc}")


def make(fasta_path: str, out_base: str, flush_bytes: int = 0):
    """FASTA -> out_base.seq/.map/.title; returns (records, slots)."""
    r, s = C.c_uint64(), C.c_uint64()
    _check(N.host_lib().swhy_dbfile_make(os.fsencode(fasta_path), os.fsencode(out_base),
                                         int(flush_bytes), C.byref(r), C.byref(s)))
    return r.value, s.value


class DbFile:
    def __init__(self, base: str):
        h = C.c_void_p()
        _check(N.host_lib().swhy_dbfile_open(os.fsencode(base), C.byref(h)))
        self.handle = h
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(N.host_lib().swhy_dbfile_info(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        self.records, self.slots, self.residues, self.batches = a.value, b.value, c.value, d.value

    def seqdb(self) -> SeqDB:
        """Every sequence (fillers as length 0), residues viewed in place."""
        r, o, l = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(N.host_lib().swhy_dbfile_arrays(self.handle, C.byref(r), C.byref(o), C.byref(l)))
        n = self.slots
        offsets = np.ctypeslib.as_array(C.cast(o, C.POINTER(C.c_uint64)), shape=(n,)).copy() \
            if n else np.zeros(0, np.uint64)
        lengths = np.ctypeslib.as_array(C.cast(l, C.POINTER(C.c_uint32)), shape=(n,)).copy() \
            if n else np.zeros(0, np.uint32)
        end = int((offsets + lengths).max()) if n else 0
        residues = np.ctypeslib.as_array(C.cast(r, C.POINTER(C.c_uint8)), shape=(max(end, 1),))
        return SeqDB(residues, offsets, lengths)

    def title(self, index: int) -> str:
        n = C.c_uint64()
        buf = C.create_string_buffer(4096)
        _check(N.host_lib().swhy_dbfile_title(self.handle, int(index), buf, 4096, C.byref(n)))
        if n.value >= 4096:
            buf = C.create_string_buffer(n.value + 1)
            _check(N.host_lib().swhy_dbfile_title(self.handle, int(index), buf, n.value + 1,
                                                  C.byref(n)))
        return buf.value.decode(errors="replace")

    def pool_search(self, query, matrix, gap_open=10, gap_extend=2, k=10, chunk_bytes=0,
                    n_workers=1, all_scores=False):
        """The worker pipeline over the file's batches in file order."""
        q = np.ascontiguousarray(query, dtype=np.uint8)
        m = np.ascontiguousarray(matrix, dtype=np.int32)
        ts = np.zeros(k, np.int32)
        ti = np.zeros(k, np.int64)
        sc = np.zeros(self.slots, np.int32) if all_scores else None
        secs, gc = C.c_double(0), C.c_double(0)
        _check(N.host_lib().swhy_dbfile_pool_search(
            self.handle, N.ptr(q), len(q), N.ptr(m), int(gap_open), int(gap_extend), int(k),
            int(chunk_bytes), int(n_workers), N.ptr(ts), N.ptr(ti),
            None if sc is None else N.ptr(sc), C.byref(secs), C.byref(gc)))
        return {"top_score": ts, "top_index": ti, "scores": sc, "seconds": secs.value,
                "gcups": gc.value}

    def close(self):
        if self.handle:
            N.host_lib().swhy_dbfile_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_fasta(path: str):
    """Every record: [(title, codes uint8)] with the reference's letter table
    (0..22, '@' = 23, letters without a code = 0xFF)."""
    L = N.host_lib()
    h = C.c_void_p()
    _check(L.swhy_fasta_open(os.fsencode(path), C.byref(h)))
    try:
        n = C.c_uint64()
        _check(L.swhy_fasta_count(h, C.byref(n)))
        out = []
        for i in range(n.value):
            codes, ln, title = C.c_void_p(), C.c_uint64(), C.c_char_p()
            _check(L.swhy_fasta_record(h, i, C.byref(codes), C.byref(ln), C.byref(title)))
            arr = (np.ctypeslib.as_array(C.cast(codes, C.POINTER(C.c_uint8)), shape=(ln.value,))
                   .copy() if ln.value else np.zeros(0, np.uint8))
            out.append((title.value.decode(errors="replace"), arr))
        return out
    finally:
        L.swhy_fasta_close(h)
