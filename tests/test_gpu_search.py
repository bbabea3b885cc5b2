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


@pytest.mark.parametrize("width", [16, 24, 32])
def test_c1_slice_all_widths(eng, width):
    db = small_db(1500, 11)
    q = G.c1_query()
    _, st = check(eng, db, q, strip_width=width)
    assert st["n_overflow"] == 0
    assert st["cells"] == db.total * len(q)


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
                                         (64, 8)])
def test_group_mode_shapes(eng, lanes, width):
    db = long_db(400 + lanes, n=40)
    q = G.queries([1234], 401)[0]
    check_db(eng, db, q, min_len=300, group_lanes=lanes, group_width=width)


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
