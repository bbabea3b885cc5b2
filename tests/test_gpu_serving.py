"""Serving loop: continuous batching on a resident database.  A result must
not depend on what else shares the device."""
import threading
import time

import numpy as np
import pytest

import oracle as O
from swhybrid_amd import datagen as G
from swhybrid_amd.engine import Engine, Server, topk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    rng = np.random.Generator(np.random.PCG64(91))
    lens = np.concatenate([G.lognormal_lengths(rng, 3000, 300, 0.7, 20, 2000),
                           G.loguniform_lengths(rng, 40, 2000, 6000)])
    db = G.SeqDB.from_lengths(rng, lens)
    eng = Engine("blosum62", 10, 2)
    d = eng.load_db(db)
    yield eng, db, d
    d.close()
    eng.close()


def queries(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [G.random_residues(rng, int(m)) for m in G.lognormal_lengths(rng, n, 300, 0.7, 30, 1500)]


def test_alone_equals_in_the_mix(setup, request):
    eng, db, d = setup
    qs = queries(12, 92)
    target = qs[0]
    want = O.search_i32(target, db, eng.matrix, 10, 2)
    srv = Server(eng, d, slices=3)
    request.addfinalizer(srv.close)
    alone = srv.wait(srv.submit(target, k=20, want_scores=True))
    assert np.array_equal(alone["scores"], want)
    # now the same query among others, submitted from several threads
    ids = {}
    def sub(i):
        ids[i] = srv.submit(qs[i], k=20, want_scores=(i == 0))
    ts = [threading.Thread(target=sub, args=(i,)) for i in range(len(qs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    res = {i: srv.wait(ids[i]) for i in ids}
    assert np.array_equal(res[0]["scores"], want)
    assert np.array_equal(res[0]["top_score"], alone["top_score"])
    assert np.array_equal(res[0]["top_index"], alone["top_index"])
    for i, q in enumerate(qs):
        s, ix = O.topk(O.search_i32(q, db, eng.matrix, 10, 2), 20)
        assert np.array_equal(res[i]["top_score"], s), i
        assert np.array_equal(res[i]["top_index"], ix), i
    st = srv.stats()
    assert st["completed"] == len(qs) + 1 and st["peak_inflight"] >= 2
    srv.close()


def request_bytes(eng, d, q, **opts):
    """Device bytes one request needs, as the server's refusal states them."""
    import re
    with Server(eng, d, memory_budget=1, **opts) as srv:
        with pytest.raises(Exception, match="memory_budget") as exc:
            srv.submit(q, k=5)
    return int(re.search(r"needs (\d+) device bytes", str(exc.value)).group(1))


def test_memory_admission_serialises(setup):
    eng, db, d = setup
    rng = np.random.Generator(np.random.PCG64(93))
    qs = [G.random_residues(rng, 400) for _ in range(5)]   # equal lengths: equal bytes
    # a budget that holds one request at a time: later ones wait for memory
    need = request_bytes(eng, d, qs[0], slices=2)
    with Server(eng, d, memory_budget=need + 1, slices=2) as srv:
        ids = [srv.submit(q, k=5) for q in qs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["peak_inflight"] == 1
    for q, r in zip(qs, res):
        s, ix = O.topk(O.search_i32(q, db, eng.matrix, 10, 2), 5)
        assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix)
        assert r["info"]["admit_ms"] >= r["info"]["submit_ms"]


def test_slices_and_forced_rescore(setup):
    eng, db, d = setup
    q = queries(1, 94)[0]
    want = O.search_i32(q, db, eng.matrix, 10, 2)
    for slices in (1, 5):
        with Server(eng, d, slices=slices, overflow_threshold=40) as srv:
            r = srv.wait(srv.submit(q, k=10, want_scores=True))
        assert np.array_equal(r["scores"], want)
        assert r["info"]["rounds"] == slices


def test_rejects_and_timeouts(setup):
    eng, db, d = setup
    with Server(eng, d, max_qlen=100, max_k=5) as srv:
        with pytest.raises(Exception, match="max_qlen"):
            srv.submit(np.zeros(101, dtype=np.uint8))
        with pytest.raises(Exception):
            srv.submit(np.zeros(10, dtype=np.uint8), k=6)
        rid = srv.submit(np.arange(20, dtype=np.uint8) % 20, k=5)
        r = srv.wait(rid, timeout_ms=60000)
    assert r["info"]["status"] == 0


def test_attached_server_refuses_other_calls(setup):
    """While a server's loop thread owns the context's stream and launch
    buffers, direct searches, rescoring, tuning and destroying the served
    database are refused; after the server is destroyed they work again."""
    _, db, _ = setup
    eng = Engine("blosum62", 10, 2)
    d = eng.load_db(db)
    q = queries(1, 95)[0]
    srv = Server(eng, d)
    for call in (lambda: eng.search(d, q), lambda: eng.search_batch(d, [q, q]),
                 lambda: eng.set_scoring("blosum62", 11, 1), lambda: eng.set_tuning([]),
                 lambda: d.close()):
        with pytest.raises(Exception, match="server"):
            call()
    with pytest.raises(Exception, match="already attached"):
        Server(eng, d)
    r = srv.wait(srv.submit(q, k=5))
    srv.close()
    got, _ = eng.search(d, q)
    s, ix = topk(got, 5)
    assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix)
    d.close()
    eng.close()


@pytest.mark.parametrize("queue", ["rounds", "live"])
def test_destroy_fails_waiting_requests_with_state(setup, queue):
    """A budget that holds one request at a time: destroy finishes the one on
    the device and fails the waiting ones with SWH_ERR_STATE, also for a
    thread already blocked waiting on one of them."""
    from swhybrid_amd import _native as N
    import ctypes as C
    # a database and queries large enough that the first request is still
    # on the device (about 0.3 s) while the waiters block and destroy runs
    rng = np.random.Generator(np.random.PCG64(98))
    db = G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, 1_000_000, 350, 0.7, 20, 2000))
    eng = Engine("blosum62", 10, 2)
    d = eng.load_db(db)
    qs = G.queries([6000] * 4, 96)
    opts = dict(slices=4, queue="rounds") if queue == "rounds" else dict(queue="live")
    need = request_bytes(eng, d, qs[0], **opts)
    srv = Server(eng, d, memory_budget=need + need // 2, **opts)
    ids = [srv.submit(q, k=5) for q in qs]
    codes = {}

    def waiter(i):
        ts = np.zeros(5, dtype=np.int32)
        ti = np.zeros(5, dtype=np.int64)
        info = N.SwhRequestInfo()
        codes[i] = N.lib().swh_server_wait(srv.handle, C.c_uint64(ids[i]), -1, N.ptr(ts),
                                           N.ptr(ti), None, C.byref(info))

    ts = [threading.Thread(target=waiter, args=(i,)) for i in range(len(ids))]
    for t in ts:
        t.start()
    time.sleep(0.05)
    srv.close()
    for t in ts:
        t.join()
    d.close()
    eng.close()
    assert codes[0] == 0
    assert set(codes.values()) <= {0, -4} and -4 in codes.values(), codes


def test_request_above_budget_refused_at_submit(setup):
    eng, db, d = setup
    with Server(eng, d, memory_budget=db.n * 4 // 2) as srv:
        with pytest.raises(Exception, match="memory_budget"):
            srv.submit(queries(1, 97)[0], k=5)


@pytest.fixture(scope="module")
def big():
    """A database of one layout (more than 1,024 sequences per CU): 300k
    sequences of log-normal length and 600 long ones (2,500-12,000)."""
    rng = np.random.Generator(np.random.PCG64(99))
    lens = np.concatenate([G.lognormal_lengths(rng, 300_000, 330, 0.7, 20, 2500),
                           G.loguniform_lengths(rng, 600, 2500, 12000)])
    db = G.SeqDB.from_lengths(rng, lens)
    eng = Engine("blosum62", 10, 2)
    d = eng.load_db(db)
    yield eng, db, d
    d.close()
    eng.close()


def test_budgeted_rounds_slice_and_stay_exact(big):
    """Budgeted rounds: with 2 ms of work per round, requests of 60-3000
    residues submitted together are cut into many slices of pair and lane
    blocks, shortest remaining work first.  Every score vector equals the
    request's search alone; the short requests finish before the long ones; a
    sample equals the oracle."""
    eng, db, d = big
    rng = np.random.Generator(np.random.PCG64(101))
    qs = G.queries([3000, 60, 900, 250, 1800, 120], 100)
    alone = [eng.search(d, q)[0].copy() for q in qs]
    with Server(eng, d, round_ms=2.0, queue="rounds") as srv:
        ids = [srv.submit(q, k=10, want_scores=True) for q in qs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    for i, (r, a) in enumerate(zip(res, alone)):
        assert np.array_equal(r["scores"], a), f"request {i} (length {len(qs[i])})"
        s, ix = topk(a, 10)
        assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix)
    # the longest request runs in several slices (each at least the round's
    # two lane items per workgroup slot: 4096 lane blocks shared by the
    # requests of the round, of this database's 2,344)
    rounds = [r["info"]["rounds"] for r in res]
    assert rounds[0] >= 3 and st["rounds"] >= max(rounds), rounds
    done = [r["info"]["done_ms"] for r in res]
    assert done[1] < done[0] and done[5] < done[0]      # 60 and 120 before 3000
    idx = np.unique(np.concatenate([np.arange(db.n - 50, db.n),
                                    rng.choice(db.n, 300, replace=False)]))
    for q, r in zip(qs[:3], res[:3]):
        want = O.search_i32(q, db.subset(idx), eng.matrix, 10, 2)
        assert np.array_equal(r["scores"][idx], want)


# ---------------------------------------------------------------------------
# Live queue: one persistent launch fed while it runs.

def test_live_queue_requests_join_the_running_launch(big):
    """A long request starts alone; shorter ones submitted while it runs are
    appended to the same launch and, being shorter, finish first.  Every
    score vector equals the request's search alone and a sample equals the
    oracle."""
    eng, db, d = big
    rng = np.random.Generator(np.random.PCG64(102))
    qs = G.queries([8000, 60, 900, 250, 1800, 120, 40, 2500], 103)
    alone = [eng.search(d, q)[0].copy() for q in qs]
    with Server(eng, d, queue="live", max_qlen=8192) as srv:
        ids = [srv.submit(qs[0], k=10, want_scores=True)]
        time.sleep(0.02)
        for q in qs[1:]:
            ids.append(srv.submit(q, k=10, want_scores=True))
            time.sleep(0.002)
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    for i, (r, a) in enumerate(zip(res, alone)):
        assert np.array_equal(r["scores"], a), f"request {i} (length {len(qs[i])})"
        s, ix = topk(a, 10)
        assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix)
        assert r["info"]["status"] == 0 and r["info"]["cells"] == len(qs[i]) * db.total
    done = [r["info"]["done_ms"] for r in res]
    admit = [r["info"]["admit_ms"] for r in res]
    # the short ones were admitted while the long one ran and finished before it
    for i in (1, 5, 6):
        assert admit[i] < done[0] and done[i] < done[0], (i, admit, done)
    assert st["completed"] == len(qs) and st["peak_inflight"] >= 2
    assert st["cells"] == db.total * sum(len(q) for q in qs)
    idx = np.unique(np.concatenate([np.arange(db.n - 50, db.n),
                                    rng.choice(db.n, 300, replace=False)]))
    for i in (1, 3, 6):
        want = O.search_i32(qs[i], db.subset(idx), eng.matrix, 10, 2)
        assert np.array_equal(res[i]["scores"][idx], want)


def test_live_queue_alone_equals_in_the_mix(setup):
    """On a small database (served on its first layout): a request alone and
    among others submitted from several threads gives the same scores."""
    eng, db, d = setup
    qs = queries(12, 104)
    target = qs[0]
    want = O.search_i32(target, db, eng.matrix, 10, 2)
    with Server(eng, d, queue="live") as srv:
        alone = srv.wait(srv.submit(target, k=20, want_scores=True))
        assert np.array_equal(alone["scores"], want)
        ids = {}

        def sub(i):
            ids[i] = srv.submit(qs[i], k=20, want_scores=True)
        ts = [threading.Thread(target=sub, args=(i,)) for i in range(len(qs))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        res = {i: srv.wait(ids[i]) for i in ids}
        st = srv.stats()
    assert np.array_equal(res[0]["scores"], want)
    for i, q in enumerate(qs):
        w = O.search_i32(q, db, eng.matrix, 10, 2)
        assert np.array_equal(res[i]["scores"], w), i
        s, ix = O.topk(w, 20)
        assert np.array_equal(res[i]["top_score"], s), i
        assert np.array_equal(res[i]["top_index"], ix), i
    assert st["completed"] == len(qs) + 1


def test_live_queue_forced_rescore(setup):
    """An overflow threshold of 40 sends most sequences to the int32 rescore,
    which runs when the launch has ended: scores stay exact."""
    eng, db, d = setup
    qs = queries(4, 105)
    with Server(eng, d, queue="live", overflow_threshold=40) as srv:
        ids = [srv.submit(q, k=10, want_scores=True) for q in qs]
        res = [srv.wait(i) for i in ids]
    for q, r in zip(qs, res):
        assert np.array_equal(r["scores"], O.search_i32(q, db, eng.matrix, 10, 2))


def test_live_queue_memory_admission_serialises(setup):
    """A budget that holds one request: the others wait for its memory and
    join the launch (or the next one) when it is freed."""
    eng, db, d = setup
    rng = np.random.Generator(np.random.PCG64(106))
    qs = [G.random_residues(rng, 400) for _ in range(5)]
    need = request_bytes(eng, d, qs[0], queue="live")
    with Server(eng, d, memory_budget=need + 1, queue="live") as srv:
        ids = [srv.submit(q, k=5) for q in qs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["peak_inflight"] == 1
    for q, r in zip(qs, res):
        s, ix = O.topk(O.search_i32(q, db, eng.matrix, 10, 2), 5)
        assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix)


def test_live_queue_arena_overflows_into_the_next_launch(setup):
    """200 requests of 8,000 residues: their profiles (387 KB each) exceed one
    launch's 64 MB arena, so later requests run in later launches whose
    arena reuses the same page-locked addresses.  Every top-k equals the
    request's search alone."""
    eng, db, d = setup
    rng = np.random.Generator(np.random.PCG64(107))
    qs = [G.random_residues(rng, 8000) for _ in range(200)]
    alone = [topk(eng.search(d, q)[0], 5) for q in qs]
    with Server(eng, d, queue="live", max_qlen=8192) as srv:
        ids = [srv.submit(q, k=5) for q in qs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["rounds"] >= 2, st
    for i, (r, (s, ix)) in enumerate(zip(res, alone)):
        assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], ix), i
