"""Writes tests/golden/dbfile/: two small FASTA files and the .seq/.map/.title
files the reference's own DBmaker (src/DBmaker.cpp, src/FastaFile.cpp,
src/Bucket.cpp, compiled into oracle/_ref/dbmaker by oracle/Makefile) makes of
them.

  clean.fa   every padded length holds a multiple of 128 sequences except the
             longest, whose rest is completed by fillers: no batch mixes
             lengths, where the reference's packing is exact
  mixed.fa   random lengths: most of the database ends in mixed-length
             batches, which the reference pads at shifted positions

Run from the repository root:  python tests/golden/make_dbfile_golden.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "dbfile")
DBMAKER = os.path.join(ROOT, "oracle", "_ref", "dbmaker")
LETTERS = "ACDEFGHIKLMNPQRSTVWYBZXacdefghiklmnpqrstvwy"
ODD = "JOU*"   # letters without a residue code


def record(rng, i, n, odd=False):
    s = "".join(rng.choice(list(LETTERS), size=n))
    if odd and n > 3:
        s = s[:2] + ODD[i % 4] + s[3:]
    lines = [s[k:k + 60] for k in range(0, len(s), 60)] or [""]
    return f">rec{i} len={n} some description\n" + "\n".join(lines) + "\n"


def clean_fasta(rng):
    lens = [20] * 256 + [37] * 128 + [40] * 128 + [9] * 384 + [61] * 50
    order = rng.permutation(len(lens))
    return "".join(record(rng, i, lens[j], odd=(i % 17 == 0)) for i, j in enumerate(order))


def mixed_fasta(rng):
    lens = list(rng.integers(1, 90, size=330)) + [24] * 140
    order = rng.permutation(len(lens))
    return "".join(record(rng, i, int(lens[j])) for i, j in enumerate(order))


def main():
    if not os.path.exists(DBMAKER):
        raise SystemExit("oracle/_ref/dbmaker missing: run `make -C oracle` where the "
                         "reference sources exist")
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(4242))
    for name, text in (("clean", clean_fasta(rng)), ("mixed", mixed_fasta(rng))):
        fa = os.path.join(OUT, f"{name}.fa")
        with open(fa, "w") as f:
            f.write(text)
        with tempfile.TemporaryDirectory() as tmp:
            subprocess.run([DBMAKER, fa, os.path.join(tmp, name)], check=True,
                           stdout=subprocess.DEVNULL)
            for ext in ("seq", "map", "title"):
                shutil.copy(os.path.join(tmp, f"{name}.{ext}"), os.path.join(OUT, f"{name}.{ext}"))
        print(name, {e: os.path.getsize(os.path.join(OUT, f"{name}.{e}")) for e in ("fa", "seq", "map", "title")})


if __name__ == "__main__":
    sys.exit(main())
