"""Seeded synthetic protein databases and queries for the five search workloads.

Conventions (fixed here and echoed into every bench record):
  * RNG: numpy PCG64.  Seeds: database 1000+c, queries 2000+c, arrivals
    3000+c, mutations 4000+c for workload c (1..5).
  * Residues: i.i.d. over the 20 standard amino acids with the Robinson &
    Robinson (1991) background frequencies, quantised to 1/65536 through a
    16-bit lookup table so multi-gigabyte databases generate quickly.
  * "log-normal mean M, sigma s": arithmetic mean M, log-space sigma s, so
    mu = ln M - s^2 / 2; lengths are rounded and clipped.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .alphabet import CODE_ORDER

ROBINSON_ORDER = "ARNDCQEGHILKMFPSTWYV"
ROBINSON_FREQ = np.array([
    .07805, .05129, .04487, .05364, .01925, .04264, .06295, .07377, .02199, .05142,
    .09019, .05744, .02243, .03856, .05203, .07120, .05841, .01330, .03216, .06441,
])


def _lut16() -> np.ndarray:
    p = ROBINSON_FREQ / ROBINSON_FREQ.sum()
    edges = np.round(np.cumsum(p) * 65536).astype(np.int64)
    edges[-1] = 65536
    lut = np.empty(65536, dtype=np.uint8)
    start = 0
    for letter, end in zip(ROBINSON_ORDER, edges):
        lut[start:end] = CODE_ORDER.index(letter)
        start = end
    return lut


_LUT = _lut16()


def random_residues(rng: np.random.Generator, n: int, chunk: int = 1 << 26) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        out[s:e] = _LUT[rng.integers(0, 65536, size=e - s, dtype=np.uint16)]
    return out


def lognormal_lengths(rng, n, mean, sigma, lo, hi) -> np.ndarray:
    mu = math.log(mean) - sigma * sigma / 2.0
    x = np.rint(rng.lognormal(mu, sigma, size=n))
    return np.clip(x, lo, hi).astype(np.uint32)


def loguniform_lengths(rng, n, lo, hi) -> np.ndarray:
    x = np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))
    return np.clip(np.rint(x), lo, hi).astype(np.uint32)


@dataclasses.dataclass
class SeqDB:
    """Concatenated residue codes with per-sequence offsets and lengths."""
    residues: np.ndarray   # uint8
    offsets: np.ndarray    # uint64, start of each sequence
    lengths: np.ndarray    # uint32

    @property
    def n(self) -> int:
        return int(self.lengths.shape[0])

    @property
    def total(self) -> int:
        return int(self.lengths.sum(dtype=np.uint64))

    def seq(self, i: int) -> np.ndarray:
        o = int(self.offsets[i])
        return self.residues[o:o + int(self.lengths[i])]

    @staticmethod
    def from_lengths(rng, lengths: np.ndarray) -> "SeqDB":
        lengths = np.asarray(lengths, dtype=np.uint32)
        offsets = np.zeros(lengths.shape[0], dtype=np.uint64)
        if lengths.shape[0] > 1:
            np.cumsum(lengths[:-1], dtype=np.uint64, out=offsets[1:])
        res = random_residues(rng, int(lengths.sum(dtype=np.uint64)))
        return SeqDB(res, offsets, lengths)

    @staticmethod
    def from_list(seqs) -> "SeqDB":
        seqs = [np.asarray(s, dtype=np.uint8) for s in seqs]
        lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
        offsets = np.zeros(len(seqs), dtype=np.uint64)
        if len(seqs) > 1:
            np.cumsum(lengths[:-1], dtype=np.uint64, out=offsets[1:])
        res = np.concatenate(seqs) if seqs else np.zeros(0, np.uint8)
        return SeqDB(res.astype(np.uint8), offsets, lengths)

    def subset(self, idx) -> "SeqDB":
        return SeqDB.from_list([self.seq(int(i)) for i in idx])


def mutate(rng, seq: np.ndarray, sub=0.08, ins=0.01, dele=0.01, p_geom=0.5) -> np.ndarray:
    """Near-copy: per position substitution / insertion / deletion, indel
    lengths geometric(p_geom).  About 90% identity at the defaults."""
    seq = np.asarray(seq, dtype=np.uint8)
    n = len(seq)
    u = rng.random(n)
    # deletions: a run of geometric length starting at each deletion event
    starts = np.flatnonzero(u < dele)
    diff = np.zeros(n + 1, dtype=np.int64)
    if starts.size:
        ends = np.minimum(n, starts + rng.geometric(p_geom, size=starts.size))
        np.add.at(diff, starts, 1)
        np.add.at(diff, ends, -1)
    keep = np.cumsum(diff[:n]) == 0
    # insertions: a geometric-length random run before the position
    ins_len = np.where((u >= dele) & (u < dele + ins),
                       rng.geometric(p_geom, size=n), 0)
    ins_len[~keep] = 0
    body = seq.copy()
    subs = rng.random(n) < sub
    if subs.any():
        body[subs] = random_residues(rng, int(subs.sum()))
    kept_vals = body[keep]
    kept_ins = ins_len[keep]
    total = kept_vals.size + int(kept_ins.sum())
    if total == 0:
        return random_residues(rng, 1)
    out = random_residues(rng, total)
    pos = np.arange(kept_vals.size) + np.cumsum(kept_ins)
    out[pos] = kept_vals
    return out


# --- the five workloads -------------------------------------------------------

def c1_db(n=100_000, seed=1001) -> SeqDB:
    """One query of length 256 against 100k sequences, lengths
    logN(mean 300, sigma 0.7) clipped to [20, 2000]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return SeqDB.from_lengths(rng, lognormal_lengths(rng, n, 300, 0.7, 20, 2000))


def c2_lengths(rng, n):
    """99.8% logN(mean 327.6, sigma 0.7) clipped [20, 35000] plus 0.2%
    log-uniform [2000, 35000]: mean about 350, heavy tail to 35,000."""
    n_tail = int(round(n * 0.002))
    body = lognormal_lengths(rng, n - n_tail, 327.6, 0.7, 20, 35000)
    tail = loguniform_lengths(rng, n_tail, 2000, 35000)
    lengths = np.concatenate([body, tail])
    rng.shuffle(lengths)
    return lengths


def c2_db(n=5_000_000, seed=1002) -> SeqDB:
    rng = np.random.Generator(np.random.PCG64(seed))
    return SeqDB.from_lengths(rng, c2_lengths(rng, n))


C2_QUERY_LENGTHS = (144, 464, 1000, 2005, 5478)


def queries(lengths, seed) -> list[np.ndarray]:
    rng = np.random.Generator(np.random.PCG64(seed))
    return [random_residues(rng, int(m)) for m in lengths]


def c1_query(seed=2001) -> np.ndarray:
    return queries([256], seed)[0]


def c2_queries(seed=2002) -> list[np.ndarray]:
    return queries(C2_QUERY_LENGTHS, seed)


def c3_queries(nq=64, seed=2003) -> list[np.ndarray]:
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = loguniform_lengths(rng, nq, 50, 5000)
    return [random_residues(rng, int(m)) for m in lens]


def c4_requests(n, qps, seed=3004, qseed=2004):
    """Poisson arrivals at `qps` and logN(mean 400, sigma 0.7) lengths clipped
    to [30, 5000].  Returns (arrival offsets in seconds, query list)."""
    rng = np.random.Generator(np.random.PCG64(seed + int(qps)))
    gaps = rng.exponential(1.0 / qps, size=n)
    arrivals = np.cumsum(gaps) - gaps[0]
    qrng = np.random.Generator(np.random.PCG64(qseed))
    lens = lognormal_lengths(qrng, n, 400, 0.7, 30, 5000)
    return arrivals, [random_residues(qrng, int(m)) for m in lens]


def c5_db(n=2_000_000, seed=1005, mseed=4005, qlens=(3000, 5478), qseed=2005):
    """90% logN(mean 350, sigma 0.7) clipped [20, 3000] plus 10% log-uniform
    [3001, 35000]; 20% of sequences replaced by near-copies of the queries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_long = int(round(n * 0.1))
    lengths = np.concatenate([
        lognormal_lengths(rng, n - n_long, 350, 0.7, 20, 3000),
        loguniform_lengths(rng, n_long, 3001, 35000),
    ])
    rng.shuffle(lengths)
    qs = queries(qlens, qseed)
    mrng = np.random.Generator(np.random.PCG64(mseed))
    planted = np.sort(mrng.choice(n, size=int(round(n * 0.2)), replace=False))
    which = mrng.integers(0, len(qs), size=planted.shape[0])
    copies = {int(p): mutate(mrng, qs[int(w)]) for p, w in zip(planted, which)}
    for p, c in copies.items():
        lengths[p] = len(c)
    db = SeqDB.from_lengths(rng, lengths)
    for p, c in copies.items():
        o = int(db.offsets[p])
        db.residues[o:o + len(c)] = c
    return db, qs, planted
