"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
kernels (avxCompute + avxCompute16 and Compute + Compute16, compiled from the
reference sources into oracle/_ref/libswref.so by oracle/Makefile).

Inputs are regenerated from seeds by swhybrid_amd.datagen; each fixture keeps
a SHA-256 of the generated residues so drift in the generator is caught.
Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from swhybrid_amd import alphabet as A  # noqa: E402
from golden_inputs import FIXTURES, FIXTURES_INT8_ONLY  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(db):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(db.lengths).tobytes())
    h.update(np.ascontiguousarray(db.residues).tobytes())
    return h.hexdigest()


def make(name, fn, int8_only=False):
    db, q, mname, go, ge = fn()
    m = A.matrix(mname)
    avx, ovf_avx, _, _ = O.ref_search(q, db, m, go, ge, path="avx2")
    sse, ovf_sse, _, _ = O.ref_search(q, db, m, go, ge, path="sse")
    arrays = dict(
        query=q, lengths=db.lengths, scores_avx2=avx, scores_sse=sse,
        overflow_avx2=ovf_avx, overflow_sse=ovf_sse,
        matrix=np.array(mname), gap_open=np.int32(go), gap_extend=np.int32(ge),
        residues_sha256=np.array(sha(db)), int8_only=np.bool_(int8_only),
    )
    if db.total <= 20000:
        arrays["residues"] = db.residues
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
    print(f"{name}: n={db.n} residues={db.total} qlen={len(q)} overflow={int(ovf_avx.sum())} "
          f"max={int(avx.max())}")


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libswref.so missing: run `make -C oracle` where the "
                         "reference sources exist")
    for k, f in FIXTURES.items():
        make(k, f)
    for k, f in FIXTURES_INT8_ONLY.items():
        make(k, f, int8_only=True)


if __name__ == "__main__":
    main()
