"""swh_search_noalloc (the entry point that allocates nothing: every device
buffer of the call is a part of the library's static scratch array) returns
exactly what swh_search returns, on two calls in a row with the second one
smaller, and refuses sizes beyond its stated limits."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from swhybrid_amd import _native as N
from swhybrid_amd import datagen as G
from swhybrid_amd.engine import Engine

pytestmark = pytest.mark.gpu

SWH_ERR_ARG, SWH_ERR_CAPACITY = -1, -5


def noalloc(eng, d, q, **opts):
    q = np.ascontiguousarray(q, dtype=np.uint8)
    out = np.full(d.n, -7, dtype=np.int32)
    st = N.SwhStats()
    o = N.SwhSearchOpts(*[int(opts.get(f, 0)) for f, _ in N.SwhSearchOpts._fields_])
    rc = N.lib().swh_search_noalloc(eng.handle, d.handle, N.ptr(q), len(q), C.byref(o),
                                    N.ptr(out), C.byref(st))
    return rc, out, st.as_dict()


def database(n, seed, long_tail=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = G.lognormal_lengths(rng, n, 300, 0.7, 20, 2000)
    if long_tail:
        lens = np.concatenate([lens, G.loguniform_lengths(rng, long_tail, 2500, 9000)])
    return G.SeqDB.from_lengths(rng, lens)


def test_two_calls_in_a_row_equal_swh_search():
    eng = Engine("blosum62", 10, 2)
    # first call: a database with long sequences (group items, borders) and a
    # query of several strips; second call: a smaller database and query
    big = database(20_000, 301, long_tail=40)
    small = database(3_000, 302)
    q1, q2 = G.queries([1500, 200], 303)
    d1 = eng.load_db(big)
    d2 = eng.load_db(small)
    want1, st1 = eng.search(d1, q1)
    want2, _ = eng.search(d2, q2)
    rc, got1, na1 = noalloc(eng, d1, q1)
    assert rc == 0, N.lib().swh_last_error()
    assert np.array_equal(got1, want1)
    assert na1["cells"] == st1["cells"] and na1["cells_group"] > 0
    rc, got2, _ = noalloc(eng, d2, q2)
    assert rc == 0, N.lib().swh_last_error()
    assert np.array_equal(got2, want2)
    # and both against the oracle on a sample (the full small one)
    assert np.array_equal(got2, O.search_i32(q2, small, eng.matrix, 10, 2))
    idx = np.unique(np.concatenate([np.arange(big.n - 40, big.n), np.arange(0, big.n, 97)]))
    assert np.array_equal(got1[idx], O.search_i32(q1, big.subset(idx), eng.matrix, 10, 2))
    d1.close()
    d2.close()
    eng.close()


def test_forced_rescore_and_every_width():
    """The int32 rescore's overflow list and the strip borders live in the
    scratch too: forced rescoring and each lane shape give swh_search's scores."""
    eng = Engine("blosum62", 10, 2)
    db = database(6_000, 304, long_tail=10)
    d = eng.load_db(db)
    q = G.queries([700], 305)[0]
    want, _ = eng.search(d, q)
    for opts in ({"overflow_threshold": 40}, {"strip_width": 16, "profile_mode": 2},
                 {"strip_width": 32, "profile_mode": 1}, {"group_lanes": 8, "group_width": 8}):
        rc, got, st = noalloc(eng, d, q, **opts)
        assert rc == 0, (opts, N.lib().swh_last_error())
        assert np.array_equal(got, want), opts
        if "overflow_threshold" in opts:
            assert st["n_overflow"] == int((want >= 40).sum()) > 0
    d.close()
    eng.close()


def test_limits_are_refused():
    eng = Engine("blosum62", 10, 2)
    db = database(500, 306)
    d = eng.load_db(db)
    rc, _, _ = noalloc(eng, d, G.queries([8193], 307)[0])
    assert rc == SWH_ERR_CAPACITY and b"8192" in N.lib().swh_last_error()
    rc, _, _ = noalloc(eng, d, G.queries([50], 308)[0], use_graph=1)
    assert rc == SWH_ERR_ARG
    # after the refusals the entry point still works
    q = G.queries([8192], 309)[0]
    rc, got, _ = noalloc(eng, d, q)
    assert rc == 0 and np.array_equal(got, eng.search(d, q)[0])
    d.close()
    eng.close()
