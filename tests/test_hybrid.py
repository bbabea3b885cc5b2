"""Hybrid pool: a CPU worker (the host's AVX2 units) beside HIPWorkers on one
DataPool.  CPU tests run the CPU worker alone (no device); GPU tests run both
kinds together.  Every score is checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from swhybrid_amd import _native as N
from swhybrid_amd import alphabet as A
from swhybrid_amd import datagen as G
from swhybrid_amd.engine import hybrid_search


def pool_order(db):
    return np.array(sorted(range(db.n), key=lambda i: (-int(db.lengths[i]), i)), dtype=np.int64)


def check(db, q, k=12, matrix="blosum62", go=10, ge=2, **kw):
    r = hybrid_search(db, q, matrix=matrix, gap_open=go, gap_extend=ge, k=k, all_scores=True, **kw)
    want = O.search_i32(q, db, A.matrix(matrix), go, ge)
    bad = np.flatnonzero(r["scores"] != want)
    assert bad.size == 0, f"{bad.size} mismatches: {r['scores'][bad[:5]]} vs {want[bad[:5]]}"
    order = pool_order(db)
    s, pidx = O.topk(want[order], k)
    assert np.array_equal(r["top_score"], s)
    assert np.array_equal(r["top_index"], np.where(s > 0, order[pidx], 0))
    return r


def mixed_db(seed, n=900, lo=1, hi=2500, mean=250):
    rng = np.random.Generator(np.random.PCG64(seed))
    return G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, n, mean, 0.8, lo, hi))


@pytest.mark.parametrize("qlen", [1, 5, 64, 333, 1500])
def test_cpu_worker_alone(qlen):
    db = mixed_db(300 + qlen)
    q = G.queries([qlen], 400 + qlen)[0]
    r = check(db, q, cpu_threads=3, hip_workers=0, cpu_chunk_bytes=50_000)
    assert r["report"]["cpu_residues"] > 0 and r["report"]["hip_residues"] == 0


def test_cpu_worker_int16_pass_and_gaps():
    """Near-copies overflow the int8 pass (scores > 244) and are finished in
    int16; gap settings include one beyond int8 (open + extend > 127)."""
    rng = np.random.Generator(np.random.PCG64(501))
    q = G.queries([700], 502)[0]
    seqs = [G.mutate(rng, q) for _ in range(40)] + [G.random_residues(rng, int(m))
                                                   for m in rng.integers(1, 900, size=300)]
    db = G.SeqDB.from_list(seqs)
    r = check(db, q, cpu_threads=2, hip_workers=0)
    assert r["top_score"][0] > 1000
    check(db, q, go=3, ge=1, cpu_threads=2, hip_workers=0)
    check(db, q, go=120, ge=10, cpu_threads=2, hip_workers=0)
    check(db, q, matrix="blosum50", go=13, ge=2, cpu_threads=1, hip_workers=0)


def test_pad_and_unknown_codes_on_cpu():
    rng = np.random.Generator(np.random.PCG64(503))
    seqs = [G.random_residues(rng, 100) for _ in range(70)]
    for s in seqs[::2]:
        s[::9] = 23
    for s in seqs[1::4]:
        s[::13] = 200            # codes above 23 count as 23
    db = G.SeqDB.from_list(seqs)
    q = G.queries([90], 504)[0]
    r = hybrid_search(db, q, k=5, hip_workers=0, cpu_threads=2, all_scores=True)
    clean = G.SeqDB.from_list([np.where(s > 23, 23, s).astype(np.uint8) for s in seqs])
    assert np.array_equal(r["scores"], O.search_i32(q, clean, A.BLOSUM62))


def test_bad_options():
    db = mixed_db(1, n=10)
    q = G.queries([20], 2)[0]
    with pytest.raises(N.SwhError):
        hybrid_search(db, q, hip_workers=0, cpu_threads=0)


def test_int16_overflow_needs_the_rescore_device():
    """Scores beyond int16 go to the RecalcWorker, which rescores on device 0;
    without a device that is an error, not a wrong score."""
    n = C.c_int(0)
    N.lib().swh_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a device is present")
    w = A.CODE_ORDER.index("W")
    db = G.SeqDB.from_list([np.full(6000, w, np.uint8), np.arange(20, dtype=np.uint8)])
    with pytest.raises(N.SwhError, match="RecalcWorker"):
        hybrid_search(db, np.full(6000, w, np.uint8), hip_workers=0, cpu_threads=1)


@pytest.mark.gpu
def test_hybrid_pool_both_kinds():
    db = mixed_db(601, n=6000, hi=4000, mean=300)
    q = G.queries([400], 602)[0]
    r = check(db, q, hip_workers=1, cpu_threads=4, chunk_bytes=200_000, cpu_chunk_bytes=40_000)
    rep = r["report"]
    assert rep["cpu_residues"] > 0 and rep["hip_residues"] > 0
    assert rep["cpu_residues"] + rep["hip_residues"] == int(sum((int(L) + 3) // 4 * 4 or 4
                                                                 for L in db.lengths))


@pytest.mark.gpu
def test_int16_overflow_rescored_exactly():
    w = A.CODE_ORDER.index("W")
    q = np.full(6000, w, np.uint8)
    rng = np.random.Generator(np.random.PCG64(603))
    seqs = [np.full(6000, w, np.uint8), np.full(5990, w, np.uint8)]
    seqs += [G.random_residues(rng, 300) for _ in range(50)]
    db = G.SeqDB.from_list(seqs)
    r = check(db, q, k=3, hip_workers=0, cpu_threads=2)
    assert r["report"]["recalc"] == 2 and r["top_score"][0] == 66000


def test_pool_prints_nothing_on_stdout(capfd):
    """Library calls keep stdout clean (bench.py prints exactly one JSON line
    there); the workers' performance lines go to stderr."""
    db = mixed_db(77, n=300)
    q = G.queries([50], 78)[0]
    hybrid_search(db, q, k=3, hip_workers=0, cpu_threads=2)
    out, err = capfd.readouterr()
    assert out == "" and "CPU[" in err
