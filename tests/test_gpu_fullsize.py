"""Parity at the benchmark's full sizes.  The 5M-sequence database and its five
query lengths are searched in full on the GPU; the oracle checks a sample
(random sequences plus the longest ones) exactly, and properties that do not
depend on size check everything: every strip width and the batched launch
give the same scores, the top-k equals a sort of the scores, and a forced
int32 rescore changes nothing.  The near-copy database (20% planted copies,
BLOSUM62 10/2 and BLOSUM50 13/2) runs at a tenth of its size, and once at full
size with BLOSUM50, where every sequence the int16 pass hands to the int32
rescore is checked."""
import numpy as np
import pytest

import oracle as O
from swhybrid_amd import alphabet as A
from swhybrid_amd import datagen as G
from swhybrid_amd.engine import Engine, topk

pytestmark = pytest.mark.gpu


def sample_idx(db, rng, n_random=1500, n_longest=100):
    longest = np.argsort(-db.lengths.astype(np.int64), kind="stable")[:n_longest]
    rand = rng.choice(db.n, size=n_random, replace=False)
    return np.unique(np.concatenate([longest, rand]))


def check_sample(q, db, got, idx, matrix, go, ge):
    want = O.search_i32(q, db.subset(idx), matrix, go, ge)
    bad = np.flatnonzero(got[idx] != want)
    assert bad.size == 0, (f"{bad.size} mismatches at {idx[bad[:5]]}: got {got[idx][bad[:5]]} "
                           f"want {want[bad[:5]]}")


def sorted_topk(scores, k):
    order = np.lexsort((np.arange(scores.size), -scores.astype(np.int64)))[:k]
    s = scores[order]
    return np.where(s > 0, s, 0), np.where(s > 0, order, 0)


@pytest.fixture(scope="module")
def c2():
    db = G.c2_db()
    e = Engine("blosum62", 10, 2)
    d = e.load_db(db)
    yield db, e, d
    d.close()
    e.close()


def test_c2_all_query_lengths(c2):
    db, e, d = c2
    rng = np.random.Generator(np.random.PCG64(7001))
    idx = sample_idx(db, rng)
    qs = G.c2_queries()
    single = []
    for q in qs:
        got, st = e.search(d, q)
        assert st["cells"] == db.total * len(q)
        check_sample(q, db, got, idx, e.matrix, 10, 2)
        s1, i1 = topk(got, 20)
        s2, i2 = sorted_topk(got, 20)
        assert np.array_equal(s1, s2) and np.array_equal(i1, i2)
        single.append(got)
    batch, _ = e.search_batch(d, qs)
    for i in range(len(qs)):
        assert np.array_equal(batch[i], single[i]), f"batched query {i}"


@pytest.mark.parametrize("width,placement", [(24, "pair"), (32, "single"), (16, "single")])
def test_c2_widths_agree(c2, width, placement):
    db, e, d = c2
    q = G.c2_queries()[2]                              # length 1000
    ref, _ = e.search(d, q, strip_width=16, profile_mode="pair")
    got, _ = e.search(d, q, strip_width=width, profile_mode=placement)
    assert np.array_equal(got, ref)


def test_c2_group_layout_agrees(c2):
    db, e, _ = c2
    q = G.c2_queries()[3]                              # length 2005
    d1 = e.load_db(db, group_min_len=600)              # many more sequences in lane groups
    d2 = e.load_db(db, group_min_len=-1)               # none
    a, st1 = e.search(d1, q)
    b, st2 = e.search(d2, q)
    d1.close()
    d2.close()
    assert st1["cells_group"] > st2["cells_group"] == 0
    assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def c5_tenth():
    return G.c5_db(n=200_000)


@pytest.mark.parametrize("matrix,go,ge", [("blosum62", 10, 2), ("blosum50", 13, 2)])
def test_near_copies_tenth_scale(c5_tenth, matrix, go, ge):
    db, qs, planted = c5_tenth
    e = Engine(matrix, go, ge)
    d = e.load_db(db)
    rng = np.random.Generator(np.random.PCG64(7002))
    idx = np.unique(np.concatenate([sample_idx(db, rng, n_random=400, n_longest=50),
                                    rng.choice(planted, size=150, replace=False)]))
    for q in qs:
        got, st = e.search(d, q)
        assert got[planted].max() > 10_000                # the copies score high
        check_sample(q, db, got, idx, e.matrix, go, ge)
        forced, st2 = e.search(d, q, overflow_threshold=5_000)
        assert st2["n_overflow"] >= int((got >= 5_000).sum()) > 0
        assert np.array_equal(forced, got)
    d.close()
    e.close()


def test_overflow_stress_full_size():
    """2M sequences (4.3G residues, 10% of 3,001-35,000 residues, 20% near-copies
    of the two queries), both queries in one launch with BLOSUM50 13/2: the
    copies of the 5478-residue query that reach the int16 pass's exact limit
    are rescored in int32; all of them, the longest sequences and a random
    sample equal the oracle."""
    db, qs, planted = G.c5_db()
    e = Engine("blosum50", 13, 2)
    d = e.load_db(db)
    got, st = e.search_batch(d, qs)
    d.close()
    e.close()
    assert st["cells"] == db.total * sum(len(q) for q in qs)
    rescored = [np.flatnonzero(got[i] >= 31744) for i in range(len(qs))]
    assert st["n_overflow"] == sum(r.size for r in rescored) > 0
    rng = np.random.Generator(np.random.PCG64(7005))
    base = sample_idx(db, rng, n_random=300, n_longest=10)
    for i, q in enumerate(qs):
        assert got[i][planted].min() > 0
        idx = np.unique(np.concatenate([base, rescored[i], planted[:20]]))
        check_sample(q, db, got[i], idx, A.matrix("blosum50"), 13, 2)


def test_c1_full_size_every_score():
    """The CPU-runnable workload in full: 100k sequences (29.9M residues), one
    query of 256, BLOSUM62 10/2.  Every score and the top-10 (score desc,
    index asc) equal the oracle."""
    db = G.c1_db()
    q = G.c1_query()
    e = Engine("blosum62", 10, 2)
    d = e.load_db(db)
    got, st = e.search(d, q)
    d.close()
    e.close()
    assert st["cells"] == db.total * len(q) and db.n == 100_000
    want = O.search_i32(q, db, A.matrix("blosum62"), 10, 2)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches at {bad[:5]}"
    s1, i1 = topk(got, 10)
    s2, i2 = O.topk(want, 10)
    assert np.array_equal(s1, s2) and np.array_equal(i1, i2)


C3_TUNE_BUCKETS = (96, 192, 384, 768, 1536, 3072, 5000)


def test_c3_batched_tuned_graph_replay(c2):
    """64 queries of log-uniform lengths in [50, 5000] over the 5M-sequence
    database in one launch, after the autotuner installed its table (and
    re-laid out the database if it chose another group threshold): the eager
    call, then its hipGraph replay, each row equal to the plain one-query
    search of the same query on an untuned engine; a sample against the
    oracle."""
    db, e0, d0 = c2
    qs = G.c3_queries()
    e = Engine("blosum62", 10, 2)
    tune = e.autotune(db, C3_TUNE_BUCKETS, sample_every=8, queries_per_bucket=4)
    assert len(e.get_tuning()) == len(C3_TUNE_BUCKETS)
    gmin = tune["group_min_len"]
    d = e.load_db(db, group_min_len=-1 if gmin >= 2**31 else gmin)
    out = e.pinned_scores(db.n, len(qs))
    _, st0 = e.search_batch(d, qs, use_graph=1, out=out)     # eager, then captured
    eager = out.copy()
    out[:] = -9
    _, st1 = e.search_batch(d, qs, use_graph=1, out=out)     # replayed
    assert st1["cells"] == st0["cells"] == db.total * sum(len(q) for q in qs)
    assert np.array_equal(out, eager), "graph replay differs from the eager call"
    del out
    d.close()
    e.close()
    row = np.empty(db.n, dtype=np.int32)
    for i, q in enumerate(qs):
        e0.search(d0, q, out=row)
        assert np.array_equal(eager[i], row), f"query {i} (length {len(q)})"
    rng = np.random.Generator(np.random.PCG64(7006))
    lens = np.array([len(q) for q in qs])
    full = set(np.argsort(lens)[[0, -1]].tolist()) | set(rng.choice(len(qs), 6, replace=False).tolist())
    for i in range(len(qs)):
        idx = (sample_idx(db, rng, n_random=1500, n_longest=20) if i in full
               else np.sort(rng.choice(db.n, size=200, replace=False)))
        check_sample(qs[i], db, eager[i], idx, A.matrix("blosum62"), 10, 2)


def test_c4_server_poisson_on_resident_5m(c2):
    """The serving loop on the resident 5M-sequence database, requests of
    log-normal length (mean 400) arriving as a Poisson process at 5 QPS: each
    request's full score vector and top-20 equal its plain search alone, and a
    sample of each equals the oracle."""
    import threading
    import time

    from swhybrid_amd.engine import Server
    db, e0, d0 = c2
    arrivals, qs = G.c4_requests(16, 5.0)
    e = Engine("blosum62", 10, 2)
    d = e.load_db(db)
    srv = Server(e, d)
    ids = [None] * len(qs)
    t0 = time.perf_counter()

    def feed():
        for i, (a, q) in enumerate(zip(arrivals, qs)):
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids[i] = srv.submit(q, k=20, want_scores=True)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    d.close()
    e.close()
    assert st["completed"] == len(qs)
    rng = np.random.Generator(np.random.PCG64(7007))
    row = np.empty(db.n, dtype=np.int32)
    for i, (q, r) in enumerate(zip(qs, res)):
        e0.search(d0, q, out=row)
        assert np.array_equal(r["scores"], row), f"request {i} differs from its search alone"
        s, ix = topk(row, 20)
        assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix), i
        check_sample(q, db, r["scores"], np.sort(rng.choice(db.n, 300, replace=False)),
                     A.matrix("blosum62"), 10, 2)
