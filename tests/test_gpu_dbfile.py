"""Searching a database in the reference's file format on the GPU: the worker
pipeline over the file's batches, and the swhybrid command line, against the
oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from swhybrid_amd import alphabet as A
from swhybrid_amd import datagen as G
from swhybrid_amd import dbfile as F
from swhybrid_amd.engine import Engine, topk

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_fasta(path, seqs, prefix):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(f">{prefix}{i} synthetic\n")
            text = A.decode(s)
            for k in range(0, len(text), 60):
                f.write(text[k:k + 60] + "\n")


@pytest.fixture(scope="module")
def filedb(tmp_path_factory):
    d = tmp_path_factory.mktemp("filedb")
    rng = np.random.Generator(np.random.PCG64(1201))
    q = G.queries([300], 1202)[0]
    seqs = [G.random_residues(rng, int(n)) for n in G.lognormal_lengths(rng, 900, 250, 0.7, 5, 2500)]
    seqs += [G.mutate(rng, q) for _ in range(12)]
    write_fasta(d / "db.fa", seqs, "s")
    F.make(str(d / "db.fa"), str(d / "db"))
    f = F.DbFile(str(d / "db"))
    yield d, f, q
    f.close()


def test_pool_search_over_file_batches(filedb):
    d, f, q = filedb
    db = f.seqdb()
    want = O.search_i32(q, db, A.matrix("blosum62"), 10, 2)
    r = f.pool_search(q, A.matrix("blosum62"), k=20, all_scores=True, chunk_bytes=64 << 10)
    assert np.array_equal(r["scores"], want)
    s, i = O.topk(want, 20)
    assert np.array_equal(r["top_score"], s) and np.array_equal(r["top_index"], i)
    assert all(f.title(int(x)).startswith(">s9") for x in r["top_index"][:12])   # the near-copies


def test_engine_on_file_arrays(filedb):
    d, f, q = filedb
    db = f.seqdb()
    e = Engine("blosum62", 10, 2)
    h = e.load_db(db)
    got, _ = e.search(h, q)
    assert np.array_equal(got, O.search_i32(q, db, e.matrix, 10, 2))
    h.close()
    e.close()


def parse_cli(text):
    blocks = re.findall(r"-+top (\d+) scores-+\n(.*?)\n-{10,}", text, flags=re.S)
    out = []
    for k, body in blocks:
        rows = [re.match(r"score (-?\d+) -- (.*)$", line) for line in body.split("\n")]
        out.append([(int(m.group(1)), m.group(2)) for m in rows])
    return out


def test_cli_single_query_worker_pipeline(filedb):
    d, f, q = filedb
    write_fasta(d / "q1.fa", [q], "query")
    out = subprocess.run([os.path.join(ROOT, "swhybrid_amd", "swhybrid"), "-db", str(d / "db"),
                          "-q", str(d / "q1.fa"), "-t", "15"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    (hits,) = parse_cli(out.stdout)
    s, i = O.topk(O.search_i32(q, f.seqdb(), A.matrix("blosum62"), 10, 2), 15)
    assert [h[0] for h in hits] == list(s)
    assert [h[1] for h in hits] == [f.title(int(x)) for x in i]
    assert "GCUPS" in out.stdout


def test_cli_query_batch(filedb):
    d, f, q = filedb
    rng = np.random.Generator(np.random.PCG64(1203))
    qs = [q, G.random_residues(rng, 40), G.random_residues(rng, 1100)]
    write_fasta(d / "qs.fa", qs, "query")
    out = subprocess.run([os.path.join(ROOT, "swhybrid_amd", "swhybrid"), "-db", str(d / "db"),
                          "-q", str(d / "qs.fa"), "-t", "5", "-o", "11", "-e", "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    blocks = parse_cli(out.stdout)
    assert len(blocks) == 3
    db = f.seqdb()
    for hits, qq in zip(blocks, qs):
        s, i = O.topk(O.search_i32(qq, db, A.matrix("blosum62"), 11, 1), 5)
        assert [h[0] for h in hits] == list(s)
        assert [h[1] for h in hits] == [f.title(int(x)) for x in i]
