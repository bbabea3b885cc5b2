"""Inputs of the golden fixtures in tests/golden/, regenerated from seeds."""
import numpy as np

from swhybrid_amd import datagen as G


def c1_mini():
    return G.c1_db(n=2000, seed=1001), G.c1_query(), "blosum62", 10, 2


def ovf_mini(qlen, n_copies, n_random, seed):
    q = G.queries([qlen], seed)[0]
    rng = np.random.Generator(np.random.PCG64(seed + 4000))
    seqs = [G.mutate(rng, q) for _ in range(n_copies)]
    seqs += [G.random_residues(rng, int(n))
             for n in G.lognormal_lengths(rng, n_random, 350, 0.7, 20, 2000)]
    return G.SeqDB.from_list(seqs), q


def tiny(seed=77, n=300):
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = rng.integers(1, 13, size=n)
    db = G.SeqDB.from_list([G.random_residues(rng, int(m)) for m in lens])
    q = G.random_residues(rng, 9)
    return db, q


FIXTURES = {
    "c1_mini": lambda: c1_mini(),
    "ovf_q464": lambda: (*ovf_mini(464, 200, 300, 5001), "blosum62", 10, 2),
    "ovf_q3000": lambda: (*ovf_mini(3000, 40, 100, 5002), "blosum62", 10, 2),
    "tiny_q9": lambda: (*tiny(), "blosum62", 10, 2),
}

# BLOSUM50 13/2 through the reference's 8-bit pass only: avxCompute takes the
# matrix and gaps as arguments, but its int16 rescore hardcodes 12/2
# (src/AVXKernel.cpp:914-915), so only lanes that did not overflow are golden.
FIXTURES_INT8_ONLY = {
    "b50_q464": lambda: (*ovf_mini(464, 20, 300, 5003), "blosum50", 13, 2),
}


def regenerate(name):
    fn = FIXTURES.get(name) or FIXTURES_INT8_ONLY[name]
    db, q, _, _, _ = fn()
    return db, q
