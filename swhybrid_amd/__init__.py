"""MI355X Smith-Waterman protein database search (HIP kernels behind a C ABI).

The compute path is libswh.so (swhybrid_amd/csrc, built for gfx950); this
package holds its Python front end, the residue alphabet and matrices, and the
synthetic workload generators.
"""
from .alphabet import BLOSUM50, BLOSUM62, encode, decode, matrix  # noqa: F401
from .datagen import SeqDB  # noqa: F401

__version__ = "0.1.0"


def engine(*args, **kwargs):
    from .engine import Engine
    return Engine(*args, **kwargs)
