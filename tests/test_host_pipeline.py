"""The C++ worker pipeline (DataPool -> HIPWorker -> RecalcWorker -> WorkerPool)
behind libswhybrid_host.so."""
import os
import re

import numpy as np
import pytest

import oracle as O
from swhybrid_amd import _native as N
from swhybrid_amd import alphabet as A
from swhybrid_amd import datagen as G
from swhybrid_amd.engine import pool_search

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_library_exports_header():
    text = open(os.path.join(ROOT, "include", "swhybrid_host.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = set(re.findall(r"\b(swhy\w+)\s*\(", text))
    assert declared == set(N.HOST_EXPORTS)
    L = N.host_lib()
    for name in declared:
        assert hasattr(L, name)


def test_pool_reports_device_failure_cleanly():
    """Without a usable device the worker's error comes back as a status."""
    import ctypes as C
    n = C.c_int(0)
    N.lib().swh_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a device is present")
    db = G.SeqDB.from_list([np.arange(20, dtype=np.uint8)])
    with pytest.raises(N.SwhError, match="swh_ctx_create"):
        pool_search(db, np.arange(10, dtype=np.uint8), k=3)


def test_pool_rejects_bad_arguments():
    L = N.host_lib()
    assert L.swhy_pool_search(None, None, None, 5, None, 0, None, 10, 2, 3, 0, 1,
                              None, None, None, None, None, None) != 0


def pool_order(db):
    """Pool order of the in-memory DataPool: longest first, ties by index."""
    return np.array(sorted(range(db.n), key=lambda i: (-int(db.lengths[i]), i)), dtype=np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("chunk_bytes,first", [(40_000, 5_000), (300_000, 20_000)])
def test_chunk_schedule_matches_oracle(chunk_bytes, first):
    """A chunk schedule (a small first chunk, each later one twice the one
    before up to chunk_bytes) scores every sequence like the oracle and runs
    the schedule it reports."""
    from swhybrid_amd.engine import hybrid_search
    rng = np.random.Generator(np.random.PCG64(73))
    q = G.queries([250], 74)[0]
    db = G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, 2500, 250, 0.7, 5, 3000))
    r = hybrid_search(db, q, k=10, hip_workers=1, cpu_threads=0, chunk_bytes=chunk_bytes,
                      first_chunk_bytes=first, all_scores=True)
    assert np.array_equal(r["scores"], O.search_i32(q, db, A.BLOSUM62))
    rep = r["report"]
    assert (rep["chunk_bytes"], rep["first_chunk_bytes"]) == (chunk_bytes, first)
    # doubling from `first`: fewer chunks than fixed `first`-sized ones, more
    # than fixed chunk_bytes ones
    assert db.total / first > rep["dev_chunks"] > db.total / chunk_bytes


@pytest.mark.gpu
@pytest.mark.parametrize("chunk_bytes", [0, 40_000, 7_000])
def test_pool_search_matches_oracle(chunk_bytes):
    rng = np.random.Generator(np.random.PCG64(71))
    q = G.queries([300], 72)[0]
    seqs = [G.mutate(rng, q) for _ in range(25)]
    seqs += [G.random_residues(rng, int(n)) for n in G.lognormal_lengths(rng, 1200, 250, 0.7, 5, 3000)]
    db = G.SeqDB.from_list(seqs)
    r = pool_search(db, q, k=15, chunk_bytes=chunk_bytes, all_scores=True)
    want = O.search_i32(q, db, A.BLOSUM62)
    assert np.array_equal(r["scores"], want)
    # ResultList semantics over the pool order (ties: first handed out wins)
    order = pool_order(db)
    s, pidx = O.topk(want[order], 15)
    assert np.array_equal(r["top_score"], s)
    assert np.array_equal(r["top_pool_index"], np.where(s > 0, pidx, 0))
    assert np.array_equal(r["top_index"], np.where(s > 0, order[pidx], 0))
    assert r["gcups"] > 0
