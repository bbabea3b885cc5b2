"""ctypes access to the CPU checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

  liboracle.so      plain-C restatement of the reference arithmetic
  _ref/libswref.so  the reference's own CPU kernels (AVX2 / SSE), present
                    only where it could be built from the reference sources
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_gotoh_i32.restype = C.c_int32
        L.orc_gotoh_i32.argtypes = [_u8p, C.c_int, _u8p, C.c_int, _i32p, C.c_int, C.c_int]
        L.orc_sat_scan.restype = C.c_int32
        L.orc_sat_scan.argtypes = [_u8p, C.c_int, _u8p, C.c_int, _i32p, C.c_int, C.c_int, C.c_int]
        L.orc_limit8.restype = C.c_int32
        L.orc_limit8.argtypes = [_i32p]
        L.orc_search.restype = None
        L.orc_search.argtypes = [_u8p, C.c_int, _u8p, _u64p, _u32p, C.c_int64, _i32p,
                                 C.c_int, C.c_int, _i32p, C.c_void_p, C.c_int]
        L.orc_search_i32.restype = None
        L.orc_search_i32.argtypes = [_u8p, C.c_int, _u8p, _u64p, _u32p, C.c_int64, _i32p,
                                     C.c_int, C.c_int, _i32p, C.c_int]
        L.orc_topk.restype = None
        L.orc_topk.argtypes = [_i32p, C.c_int64, C.c_int, _i32p, _i64p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libswref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libswref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing (built only where the reference sources exist)")
        L = C.CDLL(path)
        L.ref_search.restype = C.c_int
        L.ref_search.argtypes = [C.c_int, _u8p, C.c_int, _i32p, C.c_int, C.c_int, _u8p, _u64p,
                                 _u32p, C.c_int64, _i32p, C.c_void_p, C.c_int,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _ref = L
    return _ref


def _m(matrix):
    return np.ascontiguousarray(np.asarray(matrix, dtype=np.int32).reshape(1024))


def _q(q):
    return np.ascontiguousarray(np.asarray(q, dtype=np.uint8))


def gotoh(q, d, matrix, gap_open=10, gap_extend=2) -> int:
    q, d = _q(q), _q(d)
    return int(lib().orc_gotoh_i32(q, len(q), d, len(d), _m(matrix), gap_open, gap_extend))


def sat_scan(q, d, matrix, gap_open=10, gap_extend=2, bits=8) -> int:
    q, d = _q(q), _q(d)
    return int(lib().orc_sat_scan(q, len(q), d, len(d), _m(matrix), gap_open, gap_extend, bits))


def limit8(matrix) -> int:
    return int(lib().orc_limit8(_m(matrix)))


def _db(db):
    return (np.ascontiguousarray(db.residues, dtype=np.uint8),
            np.ascontiguousarray(db.offsets, dtype=np.uint64),
            np.ascontiguousarray(db.lengths, dtype=np.uint32))


def search(q, db, matrix, gap_open=10, gap_extend=2, nthreads=0):
    """Reference cascade (8-bit pass + int16 rescore): (scores, overflow8)."""
    q = _q(q)
    r, o, l = _db(db)
    out = np.zeros(db.n, dtype=np.int32)
    ovf = np.zeros(db.n, dtype=np.uint8)
    lib().orc_search(q, len(q), r, o, l, db.n, _m(matrix), gap_open, gap_extend, out,
                     ovf.ctypes.data, nthreads)
    return out, ovf


def search_i32(q, db, matrix, gap_open=10, gap_extend=2, nthreads=0):
    q = _q(q)
    r, o, l = _db(db)
    out = np.zeros(db.n, dtype=np.int32)
    lib().orc_search_i32(q, len(q), r, o, l, db.n, _m(matrix), gap_open, gap_extend, out,
                         nthreads)
    return out


def topk(scores, k):
    s = np.ascontiguousarray(scores, dtype=np.int32)
    os_ = np.zeros(k, dtype=np.int32)
    oi = np.zeros(k, dtype=np.int64)
    lib().orc_topk(s, s.shape[0], k, os_, oi)
    return os_, oi


def ref_search(q, db, matrix, gap_open=10, gap_extend=2, path="avx2", nthreads=0):
    """The reference's own kernels: (scores, overflow8, seconds_pass8, seconds_rescore)."""
    q = _q(q)
    r, o, l = _db(db)
    out = np.zeros(db.n, dtype=np.int32)
    ovf = np.zeros(db.n, dtype=np.uint8)
    t8, t16 = C.c_double(0), C.c_double(0)
    rc = ref().ref_search(0 if path == "avx2" else 1, q, len(q), _m(matrix), gap_open,
                          gap_extend, r, o, l, db.n, out, ovf.ctypes.data, nthreads,
                          C.byref(t8), C.byref(t16))
    if rc != 0:
        raise RuntimeError(f"ref_search failed ({rc})")
    return out, ovf, t8.value, t16.value
