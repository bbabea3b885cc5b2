"""GPU parity: libswh.so scores against the oracle, bit for bit.

Every test goes through the C ABI (ctypes) onto cuda:0.  Sizes keep the CPU
oracle within a few seconds.
"""
import numpy as np
import pytest

import oracle as O
from swhybrid_amd import alphabet as A
from swhybrid_amd import datagen as G
from swhybrid_amd.engine import Engine, topk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine("blosum62", 10, 2)
    yield e
    e.close()


def small_db(n, seed, mean=300, lo=20, hi=2000):
    rng = np.random.Generator(np.random.PCG64(seed))
    return G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, n, mean, 0.7, lo, hi))


def check(eng, db, q, **opts):
    d = eng.load_db(db)
    got, st = eng.search(d, q, **opts)
    want = O.search_i32(q, db, eng.matrix, eng.gap_open, eng.gap_extend)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (f"{bad.size} mismatches, first {bad[:5]}: got {got[bad[:5]]} "
                           f"want {want[bad[:5]]}")
    d.close()
    return got, st


@pytest.mark.parametrize("width,placement", [(16, "pair"), (24, "pair"), (32, "single"),
                                             (16, "single"), (24, "single")])
def test_c1_slice_all_widths(eng, width, placement):
    db = small_db(1500, 11)
    q = G.c1_query()
    _, st = check(eng, db, q, strip_width=width, profile_mode=placement)
    assert st["n_overflow"] == 0
    assert st["cells"] == db.total * len(q)


@pytest.mark.parametrize("qlen", [5, 1345, 2700])
def test_single_profile_long_and_short_queries(eng, qlen):
    """Single-profile placement at W=16/24 crosses the 1344-column staging
    segments for long queries."""
    db = small_db(300, 500 + qlen, mean=250, hi=900)
    q = G.queries([qlen], 600 + qlen)[0]
    for w in (16, 24):
        check(eng, db, q, strip_width=w, profile_mode="single")


def test_pair_table_rejected_at_width_32(eng):
    db = small_db(10, 3)
    d = eng.load_db(db)
    with pytest.raises(Exception, match="pair"):
        eng.search(d, G.queries([20], 1)[0], strip_width=32, profile_mode="pair")
    d.close()


@pytest.mark.parametrize("qlen", [1, 2, 7, 33, 100, 1343, 1344, 1345, 2700])
def test_query_lengths(eng, qlen):
    db = small_db(300, 100 + qlen, mean=200, hi=800)
    q = G.queries([qlen], 7 + qlen)[0]
    check(eng, db, q)


def test_ragged_and_tiny_sequences(eng):
    rng = np.random.Generator(np.random.PCG64(5))
    lens = [1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 63, 64, 65, 127, 128, 129, 300, 1, 1]
    seqs = [G.random_residues(rng, n) for n in lens]
    db = G.SeqDB.from_list(seqs * 9)          # 180 sequences: a partial last block
    check(eng, db, G.queries([77], 3)[0])


def test_single_sequence_and_empty(eng):
    rng = np.random.Generator(np.random.PCG64(6))
    db = G.SeqDB.from_list([G.random_residues(rng, 50)])
    check(eng, db, G.queries([40], 4)[0])
    empty = G.SeqDB.from_list([])
    d = eng.load_db(empty)
    got, st = eng.search(d, G.queries([40], 4)[0])
    assert got.shape == (0,) and st["cells"] == 0


def test_zero_length_sequences_among_others(eng):
    """Zero-length sequences anywhere in the database (first, between others,
    last) score 0; the rest score as the oracle's, with an odd number of long
    sequences (a pair block with one empty slot)."""
    rng = np.random.Generator(np.random.PCG64(108))
    seqs = [np.zeros(0, np.uint8)]
    for k in range(301):
        seqs.append(G.random_residues(rng, int(rng.integers(20, 400))))
        if k % 50 == 0:
            seqs.append(np.zeros(0, np.uint8))
    seqs += [G.random_residues(rng, 2100 + 7 * k) for k in range(3)]   # three long ones
    seqs.append(np.zeros(0, np.uint8))
    db = G.SeqDB.from_list(seqs)
    q = G.queries([150], 109)[0]
    for gmin in (0, 2048):
        d = eng.load_db(db, group_min_len=gmin)
        try:
            got, _ = eng.search(d, q)
            want = O.search_i32(q, db, eng.matrix, eng.gap_open, eng.gap_extend)
            assert np.array_equal(got, want)
            assert (got[db.lengths == 0] == 0).all()
        finally:
            d.close()


def test_pad_and_unknown_codes(eng):
    rng = np.random.Generator(np.random.PCG64(8))
    seqs = [G.random_residues(rng, 120) for _ in range(130)]
    for s in seqs[::3]:
        s[::7] = 23                            # pad code inside a sequence
    for s in seqs[1::3]:
        s[::5] = 20                            # X
    db = G.SeqDB.from_list(seqs)
    q = G.queries([90], 9)[0]
    q[::11] = 23
    check(eng, db, q)
    # codes above 23 in the database behave as 23 (src/Worker.cpp:150-154)
    hi = G.SeqDB.from_list([np.where(s == 23, 200, s).astype(np.uint8) for s in seqs])
    d = eng.load_db(hi)
    got, _ = eng.search(d, q)
    want = O.search_i32(q, db, eng.matrix, 10, 2)
    assert (got == want).all()


def test_forced_int32_rescore_matches(eng):
    db = small_db(400, 21)
    q = G.queries([200], 22)[0]
    got, st = check(eng, db, q, overflow_threshold=1)
    assert st["n_overflow"] == int((got >= 1).sum())


def test_near_copies_blosum50(eng):
    rng = np.random.Generator(np.random.PCG64(31))
    q = G.queries([1500], 32)[0]
    seqs = [G.mutate(rng, q) for _ in range(40)] + [G.random_residues(rng, 400) for _ in range(100)]
    db = G.SeqDB.from_list(seqs)
    e2 = Engine("blosum50", 13, 2)
    got, st = check(e2, db, q)
    assert got[:40].min() > 1000
    got2, _ = check(e2, db, q, overflow_threshold=2000)
    e2.close()


def test_topk_matches_resultlist(eng):
    db = small_db(2000, 41)
    q = G.c1_query()
    d = eng.load_db(db)
    got, _ = eng.search(d, q)
    s1, i1 = topk(got, 20)
    s2, i2 = O.topk(O.search_i32(q, db, eng.matrix), 20)
    assert (s1 == s2).all() and (i1 == i2).all()


def test_reference_cascade_equals_gpu(eng):
    """The reference-restating cascade (8-bit pass + int16 rescore) agrees too."""
    rng = np.random.Generator(np.random.PCG64(51))
    q = G.queries([464], 52)[0]
    seqs = [G.mutate(rng, q) for _ in range(30)] + [G.random_residues(rng, 350) for _ in range(300)]
    db = G.SeqDB.from_list(seqs)
    d = eng.load_db(db)
    got, _ = eng.search(d, q)
    casc, ovf = O.search(q, db, eng.matrix, 10, 2)
    assert ovf.sum() >= 30
    assert (got == casc).all()


# --- group mode (long sequences, lanes cooperating on a pair) ------------------

def long_db(seed, n=60, lo=300, hi=5000):
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = G.loguniform_lengths(rng, n, lo, hi)
    return G.SeqDB.from_lengths(rng, lens)


def check_db(eng, db, q, min_len, **opts):
    d = eng.load_db(db, group_min_len=min_len)
    got, st = eng.search(d, q, **opts)
    want = O.search_i32(q, db, eng.matrix, eng.gap_open, eng.gap_extend)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (f"{bad.size} mismatches (opts {opts}), first {bad[:5]}: "
                           f"got {got[bad[:5]]} want {want[bad[:5]]} len {db.lengths[bad[:5]]}")
    d.close()
    return got, st


@pytest.mark.parametrize("qlen", [1, 50, 144, 700, 1100, 2700])
def test_group_mode_auto(eng, qlen):
    db = long_db(200 + qlen)
    q = G.queries([qlen], 300 + qlen)[0]
    check_db(eng, db, q, min_len=600)


@pytest.mark.parametrize("lanes,width", [(1, 8), (3, 16), (7, 8), (16, 16), (63, 8), (64, 16),
                                         (64, 8), (5, 24), (21, 24), (64, 24)])
def test_group_mode_shapes(eng, lanes, width):
    db = long_db(400 + lanes, n=40)
    q = G.queries([1234], 401)[0]
    check_db(eng, db, q, min_len=300, group_lanes=lanes, group_width=width)


@pytest.mark.parametrize("strip,placement,width", [
    (16, "pair", 8), (16, "pair", 16), (16, "single", 8), (16, "single", 16), (24, "pair", 8),
    (24, "single", 16), (24, "pair", 24), (24, "single", 24), (32, "single", 8),
    (32, "single", 16)])
def test_group_steps_every_kernel(eng, strip, placement, width):
    """Group items in every scan kernel: one database row per step where the
    kernel's register budget allows only that (16-column lanes with 16-column
    groups), two elsewhere; 3 lanes per group make several passes, so the
    pass borders of both rows of a step are written and read back."""
    db = long_db(700 + strip + width, n=30)
    q = G.queries([500], 701)[0]
    check_db(eng, db, q, min_len=300, strip_width=strip, profile_mode=placement,
             group_width=width, group_lanes=3)


def test_group_mode_odd_count_and_overflow(eng):
    rng = np.random.Generator(np.random.PCG64(61))
    q = G.queries([900], 62)[0]
    seqs = [G.mutate(rng, q) for _ in range(7)] + [G.random_residues(rng, int(n))
                                                  for n in rng.integers(10, 3000, size=91)]
    db = G.SeqDB.from_list(seqs)
    check_db(eng, db, q, min_len=1000)                        # odd number of long sequences
    got, st = check_db(eng, db, q, min_len=1000, overflow_threshold=500)
    assert st["n_overflow"] >= 7
    check_db(eng, db, q, min_len=1)                           # everything in pairs


# --- many queries in one launch ---------------------------------------------------

def test_batch_equals_single_and_oracle(eng):
    rng = np.random.Generator(np.random.PCG64(81))
    db = long_db(82, n=150, lo=20, hi=4000)
    qs = [G.random_residues(rng, int(m)) for m in (1, 30, 144, 333, 1000, 1500, 17, 2048)]
    qs.append(G.mutate(rng, db.seq(3)))                      # one strong hit
    d = eng.load_db(db, group_min_len=1200)
    got, st = eng.search_batch(d, qs)
    assert got.shape == (len(qs), db.n)
    for i, q in enumerate(qs):
        want = O.search_i32(q, db, eng.matrix, 10, 2)
        assert np.array_equal(got[i], want), f"query {i} (len {len(q)})"
        single, _ = eng.search(d, q)
        assert np.array_equal(single, want)
    assert st["cells"] == sum(db.total * len(q) for q in qs)
    forced, st2 = eng.search_batch(d, qs, overflow_threshold=200)
    assert st2["n_overflow"] > 0
    assert np.array_equal(forced, got)
    d.close()


def test_batch_empty_and_one(eng):
    db = small_db(50, 83)
    d = eng.load_db(db)
    got, st = eng.search_batch(d, [])
    assert got.shape == (0, db.n)
    q = G.queries([64], 84)[0]
    got, _ = eng.search_batch(d, [q])
    assert np.array_equal(got[0], O.search_i32(q, db, eng.matrix, 10, 2))


# --- captured launch sequences --------------------------------------------------

def test_graph_replays_match_eager(eng):
    """use_graph: the first call of a shape runs eagerly and captures; later calls
    with the same lengths replay.  Eager searches of other shapes run between
    replays (reusing the staging buffers); every replay is checked."""
    rng = np.random.Generator(np.random.PCG64(91))
    db = long_db(92, n=120, lo=20, hi=3000)
    d = eng.load_db(db, group_min_len=1500)
    lens = (40, 500, 1300)
    for rep in range(4):
        qs = [G.random_residues(rng, m) for m in lens]
        if rep == 2:
            qs[1] = G.mutate(rng, db.seq(5))[:500]          # a strong hit
            qs[1] = np.concatenate([qs[1], G.random_residues(rng, 500 - len(qs[1]))])
        got, st = eng.search_batch(d, qs, use_graph=1)
        for i, q in enumerate(qs):
            want = O.search_i32(q, db, eng.matrix, 10, 2)
            assert np.array_equal(got[i], want), f"replay {rep} query {i}"
        assert st["cells"] == sum(db.total * len(q) for q in qs)
        # other work between replays: a different shape and a single search
        other = [G.random_residues(rng, 77), G.random_residues(rng, 2100)]
        g2, _ = eng.search_batch(d, other)
        assert np.array_equal(g2[0], O.search_i32(other[0], db, eng.matrix, 10, 2))
        s1, _ = eng.search(d, other[1], strip_width=32)
        assert np.array_equal(s1, g2[1])
    # forced rescore through a captured graph
    qs = [G.random_residues(rng, m) for m in lens]
    for rep in range(2):
        got, st = eng.search_batch(d, qs, use_graph=1, overflow_threshold=30)
        want = [O.search_i32(q, db, eng.matrix, 10, 2) for q in qs]
        n_ovf = sum(int((w >= 30).sum()) for w in want)
        assert n_ovf > 0 and st["n_overflow"] == n_ovf
        for i in range(len(qs)):
            assert np.array_equal(got[i], want[i]), f"forced rescore, replay {rep} query {i}"
    d.close()


def test_graph_not_replayed_after_set_scoring():
    """A captured graph holds the gaps and the overflow threshold in its kernel
    arguments: after set_scoring the same call must not replay it.  Capture
    at 10/2, switch to 11/1 (and then to BLOSUM50 13/2), call again with the
    same lengths: the scores follow the new scoring."""
    e = Engine("blosum62", 10, 2)
    rng = np.random.Generator(np.random.PCG64(93))
    db = long_db(94, n=150, lo=20, hi=2500)
    d = e.load_db(db, group_min_len=1200)
    qs = [G.random_residues(rng, m) for m in (60, 700)]
    for rep in range(2):                           # eager + capture, then a replay
        got, _ = e.search_batch(d, qs, use_graph=1)
        for i, q in enumerate(qs):
            assert np.array_equal(got[i], O.search_i32(q, db, e.matrix, 10, 2)), (rep, i)
    for matrix, go, ge in (("blosum62", 11, 1), ("blosum50", 13, 2), ("blosum62", 10, 2)):
        e.set_scoring(matrix, go, ge)
        for rep in range(2):
            got, _ = e.search_batch(d, qs, use_graph=1)
            for i, q in enumerate(qs):
                want = O.search_i32(q, db, A.matrix(matrix), go, ge)
                assert np.array_equal(got[i], want), (matrix, go, ge, rep, i)
    d.close()
    e.close()


# --- autotuner ---------------------------------------------------------------------

def test_autotune_installs_a_table_and_scores_stay_exact():
    e = Engine("blosum62", 10, 2)
    rng = np.random.Generator(np.random.PCG64(101))
    lens = np.concatenate([G.lognormal_lengths(rng, 3000, 300, 0.7, 20, 2000),
                           G.loguniform_lengths(rng, 40, 2000, 6000)])
    db = G.SeqDB.from_lengths(rng, lens)
    res = e.autotune(db, [64, 400, 1500], sample_every=2, queries_per_bucket=2)
    assert res["sample_seqs"] == (db.n + 1) // 2
    assert res["group_min_len"] in (1024, 2048, 4096, 8192, 0xFFFFFFFF)
    ents = res["entries"]
    assert [t["qlen_max"] for t in ents] == [64, 400, 1500]
    for t in ents:
        assert t["strip_width"] in (16, 24, 32) and t["profile_mode"] in (1, 2)
        assert t["group_width"] in (0, 8, 16, 24) and 0 <= t["group_lanes"] <= 64
        assert t["group_width"] != 0 or t["group_lanes"] == 0
        assert t["group_width"] != 24 or t["strip_width"] == 24
        assert all(c > 0 for c in t["lane_cost"])
    assert e.get_tuning() == ents
    d = e.load_db(db, group_min_len=res["group_min_len"] if res["group_min_len"] < 2**31 else -1)
    qs = [G.random_residues(rng, m) for m in (30, 300, 1200, 3000)]
    got, _ = e.search_batch(d, qs)
    for i, q in enumerate(qs):
        assert np.array_equal(got[i], O.search_i32(q, db, e.matrix, 10, 2)), f"query {i}"
    d.close()
    e.close()


@pytest.mark.parametrize("cfg", range(5))
def test_tuning_table_drives_the_shape(cfg):
    """A table whose only measured lane configuration is cfg forces it; group
    lanes and widths come from the buckets.  Scores equal the oracle's."""
    from swhybrid_amd import _native as N
    e = Engine("blosum62", 10, 2)
    w, pair = N.LANE_CONFIGS[cfg]
    cost = [0.0] * 5
    cost[cfg] = 1.0
    table = [dict(qlen_max=200, strip_width=w, profile_mode=2 if pair else 1, group_lanes=3,
                  group_width=8, lane_cost=cost),
             dict(qlen_max=2000, strip_width=w, profile_mode=2 if pair else 1, group_lanes=0,
                  group_width=16, lane_cost=cost)]
    e.set_tuning(table)
    assert e.get_tuning() == table
    db = long_db(700 + cfg, n=80, lo=20, hi=3000)
    d = e.load_db(db, group_min_len=1000)
    rng = np.random.Generator(np.random.PCG64(701))
    qs = [G.random_residues(rng, m) for m in (50, 150, 900, 2500)]
    got, _ = e.search_batch(d, qs)
    for i, q in enumerate(qs):
        assert np.array_equal(got[i], O.search_i32(q, db, e.matrix, 10, 2)), f"query {i}"
    e.set_tuning([])
    assert e.get_tuning() == []
    d.close()
    e.close()


def test_tuning_table_validation():
    e = Engine("blosum62", 10, 2)
    good = dict(qlen_max=100, strip_width=16, profile_mode=2, group_lanes=0, group_width=16,
                lane_cost=[1, 1, 1, 1, 1])
    for bad in (dict(strip_width=20), dict(profile_mode=0), dict(group_width=12),
                dict(group_width=24),                      # 24 needs 24-column lanes
                dict(strip_width=32, profile_mode=2), dict(lane_cost=[0, 0, 0, 0, 0]),
                dict(group_lanes=65), dict(group_width=0, group_lanes=3)):
        with pytest.raises(Exception):
            e.set_tuning([{**good, **bad}])
    with pytest.raises(Exception):
        e.set_tuning([good, {**good, "qlen_max": 50}])     # not ascending
    e.set_tuning([good, {**good, "qlen_max": 200, "strip_width": 24, "group_width": 24}])
    e.set_tuning([good, {**good, "qlen_max": 200, "group_width": 0}])   # built-in group mode
    e.set_tuning([good])
    e.close()


@pytest.mark.parametrize("qlen", [20, 50, 90])
def test_short_query_over_long_sequences_default_shape(eng, qlen):
    """A short query over a small database of long sequences: the default
    launch shape (group width 8 where the longest pair bounds the launch)
    scores like the oracle."""
    rng = np.random.Generator(np.random.PCG64(103 + qlen))
    db = G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, 3000, 500, 0.8, 20, 2000))
    q = G.queries([qlen], 104 + qlen)[0]
    d = eng.load_db(db)
    try:
        got, st = eng.search(d, q)
        assert st["cells_group"] > 0
        assert np.array_equal(got, O.search_i32(q, db, eng.matrix, eng.gap_open, eng.gap_extend))
    finally:
        d.close()


def test_sequences_scattered_over_a_sparse_buffer(eng):
    """Sequences whose span in the caller's buffer is mostly gaps (here 80 MB
    for 0.6 MB of residues) go up as a compact copy; scores equal the oracle's."""
    rng = np.random.Generator(np.random.PCG64(105))
    dense = small_db(2000, 106)
    buf = np.full(80 << 20, 23, dtype=np.uint8)
    starts = np.sort(rng.choice((80 << 20) // 4096 - 1, dense.n, replace=False)) * 4096
    for i in range(dense.n):
        buf[starts[i]:starts[i] + dense.lengths[i]] = dense.seq(i)
    sparse = G.SeqDB(buf, starts.astype(np.uint64), dense.lengths.copy())
    q = G.queries([300], 107)[0]
    d = eng.load_db(sparse)
    try:
        assert np.array_equal(eng.search(d, q)[0],
                              O.search_i32(q, dense, eng.matrix, eng.gap_open, eng.gap_extend))
    finally:
        d.close()


def test_context_destroy_refused_while_its_databases_exist():
    """swh_ctx_destroy refuses while a database made on the context exists
    (its buffers go back to the context when it is destroyed); Engine.close
    closes its databases first."""
    from swhybrid_amd import _native as N
    e = Engine("blosum62", 10, 2)
    db = small_db(300, 97)
    d = e.load_db(db)
    d2 = e.load_db(db)
    assert N.lib().swh_ctx_destroy(e.handle) == -4            # SWH_ERR_STATE
    assert b"still exist" in N.lib().swh_last_error()
    q = G.queries([120], 98)[0]
    assert np.array_equal(e.search(d, q)[0], O.search_i32(q, db, e.matrix, 10, 2))
    d2.close()
    e.close()
    assert d.handle is None and e.handle is None


def test_build_beside_a_running_search_and_pageable_uploads(eng):
    """A database built on one thread while a search runs on another (the
    worker pipeline's overlap) scores like the oracle; so do databases whose
    residues come from pageable memory, small (through the page-locked
    staging) and above 64 MB (a prefix the runtime stages and a padded tail),
    and one from page-locked memory."""
    import threading
    from swhybrid_amd import _native as N
    big = small_db(120_000, 99, mean=600, lo=20, hi=3000)        # ~72M residues
    assert big.total > (64 << 20)
    q = G.queries([2000], 100)[0]
    d_big = eng.load_db(big)
    got = {}

    def search():
        got["big"] = eng.search(d_big, q)[0]

    small = small_db(5000, 101)
    t = threading.Thread(target=search)
    t.start()
    d_small = eng.load_db(small)                                  # beside the search
    t.join()
    assert np.array_equal(eng.search(d_small, q)[0],
                          O.search_i32(q, small, eng.matrix, eng.gap_open, eng.gap_extend))
    idx = np.random.Generator(np.random.PCG64(102)).choice(big.n, 400, replace=False)
    want = O.search_i32(q, big.subset(idx), eng.matrix, eng.gap_open, eng.gap_extend)
    assert np.array_equal(got["big"][idx], want)
    # page-locked residues go up as they are
    pin = N.pinned_array(small.total, np.uint8)
    pin[:] = small.residues
    pinned = G.SeqDB(pin, small.offsets, small.lengths)
    d_pin = eng.load_db(pinned)
    assert np.array_equal(eng.search(d_pin, q)[0], eng.search(d_small, q)[0])
    for x in (d_big, d_small, d_pin):
        x.close()


def test_small_database_second_layout(eng):
    """A database too small to fill the device with lane items of one query
    carries a second layout with every sequence of 128+ residues in lane
    groups: few-query launches run on it, launches with many lane items on the
    lane layout, and both give the oracle's scores."""
    rng = np.random.Generator(np.random.PCG64(91))
    db = G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, 4000, 300, 0.7, 20, 1900))
    d = eng.load_db(db)                                   # default layout choice
    q = G.queries([300], 92)[0]
    one, st = eng.search(d, q)
    assert st["cells_group"] > 0                          # no sequence reaches 2048
    assert np.array_equal(one, O.search_i32(q, db, eng.matrix, eng.gap_open, eng.gap_extend))
    qs = [G.random_residues(rng, 24) for _ in range(260)]  # 260 x 8 lane items
    many, stb = eng.search_batch(d, qs)
    assert stb["cells_group"] == 0
    for i in (0, 77, 259):
        assert np.array_equal(many[i], eng.search(d, qs[i])[0])
        assert np.array_equal(many[i], O.search_i32(qs[i], db, eng.matrix, eng.gap_open,
                                                     eng.gap_extend))
    single = eng.load_db(db, group_min_len=2048)          # explicit: one layout
    assert single.info()["device_bytes"] < d.info()["device_bytes"]
    _, st1 = eng.search(single, q)
    assert st1["cells_group"] == 0
    single.close()
    one_q = eng.load_db(db, queries_per_launch=1)         # only the layout one query uses
    assert one_q.info()["device_bytes"] < d.info()["device_bytes"]
    got, st2 = eng.search(one_q, q)
    assert st2["cells_group"] == st["cells_group"] and np.array_equal(got, one)
    one_q.close()
    d.close()


def test_page_locked_output(eng):
    """Scores copied by the device straight into a page-locked caller buffer
    equal those through the staging copy, eager and replayed from a graph,
    single and batched; a reused buffer is overwritten each call."""
    rng = np.random.Generator(np.random.PCG64(93))
    db = G.SeqDB.from_lengths(rng, G.lognormal_lengths(rng, 3000, 300, 0.8, 10, 3000))
    d = eng.load_db(db, group_min_len=1500)
    qs = [G.random_residues(rng, m) for m in (77, 400, 1300)]
    pin = eng.pinned_scores(db.n)
    for g in (0, 1, 1, 1):
        for q in qs:
            ref, _ = eng.search(d, q, use_graph=g)
            got, _ = eng.search(d, q, out=pin, use_graph=g)
            assert got is pin and np.array_equal(got, ref)
    many = eng.pinned_scores(db.n, nq=len(qs))
    ref, _ = eng.search_batch(d, qs)
    got, _ = eng.search_batch(d, qs, out=many)
    assert np.array_equal(got, ref)
    assert np.array_equal(ref[1], O.search_i32(qs[1], db, eng.matrix, eng.gap_open, eng.gap_extend))
    with pytest.raises(ValueError):
        eng.search(d, qs[0], out=np.empty(db.n - 1, dtype=np.int32))
    d.close()


def test_group_width_24_needs_24_column_lanes(eng):
    db = long_db(77, n=20)
    d = eng.load_db(db, group_min_len=300)
    q = G.queries([500], 78)[0]
    with pytest.raises(Exception):
        eng.search(d, q, group_width=24, strip_width=16)
    a, st = eng.search(d, q, group_width=24)
    b, _ = eng.search(d, q, group_width=16)
    assert st["cells_group"] > 0 and np.array_equal(a, b)
    d.close()


# --- extremes ------------------------------------------------------------------------

def test_very_long_query(eng):
    db = small_db(120, 901, mean=300, hi=1500)
    q = G.queries([9000], 902)[0]
    for w in (16, 24, 32):
        check(eng, db, q, strip_width=w)


def test_very_long_database_sequences(eng):
    rng = np.random.Generator(np.random.PCG64(903))
    q = G.queries([300], 904)[0]
    seqs = [G.random_residues(rng, 60_000), G.random_residues(rng, 45_001),
            G.mutate(rng, np.concatenate([G.random_residues(rng, 30_000), q]))]
    seqs += [G.random_residues(rng, int(n)) for n in rng.integers(1, 500, size=40)]
    db = G.SeqDB.from_list(seqs)
    check_db(eng, db, q, min_len=2048)
    check_db(eng, db, q, min_len=2048, group_lanes=4)
    check_db(eng, db, q, min_len=-1)                  # the same sequences on single lanes


def test_many_queries_one_launch(eng):
    rng = np.random.Generator(np.random.PCG64(905))
    db = small_db(200, 906, mean=150, hi=600)
    qs = [G.random_residues(rng, int(m)) for m in rng.integers(1, 200, size=300)]
    d = eng.load_db(db)
    got, st = eng.search_batch(d, qs)
    for i in range(0, len(qs), 7):
        assert np.array_equal(got[i], O.search_i32(qs[i], db, eng.matrix, 10, 2)), f"query {i}"
    assert st["cells"] == sum(db.total * len(q) for q in qs)
    d.close()


def test_database_buffers_reused():
    """A context keeps a destroyed database's device buffers for its next
    database: a smaller database built into a larger one's buffers (stale
    words beyond its own blocks), then a larger one again, score exactly."""
    e = Engine("blosum62", 10, 2)
    q = G.queries([400], 907)[0]
    try:
        for n, seed, hi in ((3000, 908, 3000), (700, 909, 900), (2500, 910, 2500)):
            db = small_db(n, seed, mean=300, hi=hi)
            d = e.load_db(db)
            got, _ = e.search(d, q)
            assert np.array_equal(got, O.search_i32(q, db, e.matrix, 10, 2)), (n, seed)
            d.close()
    finally:
        e.close()
