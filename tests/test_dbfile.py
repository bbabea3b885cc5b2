"""The reference's database files (.seq/.map/.title), the DBmaker and FASTA
input: CPU tests of libswhybrid_host.so against the reference DBmaker's own
output (tests/golden/dbfile, written by tests/golden/make_dbfile_golden.py)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from swhybrid_amd import dbfile as F

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dbfile")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# the reference's letter table (src/DataPool.cpp:33-47), restated for checking
_ORDER = "ABCDEFGHIKLMNPQRSTVWXYZ"


def ref_code(ch):
    if ch == "@":
        return 23
    u = ch.upper()
    return _ORDER.index(u) if u in _ORDER else 0xFF


def parse_fasta(path):
    recs, title = [], None
    with open(path) as f:
        for line in f:
            line = line.rstrip("\r\n")
            if line.startswith(">"):
                title = line
                recs.append([title, ""])
            elif title is not None:
                recs[-1][1] += line.replace(" ", "").replace("\t", "")
    return recs


def padded_codes(seq):
    n = max(4, (len(seq) + 3) // 4 * 4)
    return bytes([ref_code(c) for c in seq] + [23] * (n - len(seq)))


def seq_batches(path):
    d = open(path, "rb").read()
    end, zero = struct.unpack("<QQ", d[:16])
    assert end == len(d) and zero == 0
    p, out = 16, []
    while p < end:
        assert d[p:p + 1] == b"$" and d[p + 17:p + 18] == b"#"
        n, L = struct.unpack("<QQ", d[p + 1:p + 17])
        p += 18
        out.append((n, L, [d[p + k * L:p + (k + 1) * L] for k in range(n)]))
        p += n * L
    return out


def check_db(base, fasta, exact_sequences=True):
    """Every record once, in title order, with its padded codes."""
    recs = {t: s for t, s in parse_fasta(fasta)}
    titles = open(base + ".title").read().split("\n")[:-1]
    assert sorted(titles) == sorted(recs)
    batches = seq_batches(base + ".seq")
    slots = [s for _, _, ss in batches for s in ss]
    assert len(slots) % 128 == 0 and len(slots) - len(titles) < 128
    for s in slots[len(titles):]:
        assert set(s) == {23}                                  # fillers
    if exact_sequences:
        for i, t in enumerate(titles):
            want = padded_codes(recs[t])
            assert slots[i][:len(want)] == want and set(slots[i][len(want):]) <= {23}, t
    # map: record count, then per 128 slots the offset of its first title
    m = np.fromfile(base + ".map", dtype="<u8")
    assert m[0] == len(titles)
    text = open(base + ".title", "rb").read()
    starts = [0]
    for i, ch in enumerate(text):
        if ch == 10:
            starts.append(i + 1)
    assert list(m[1:]) == [starts[g] for g in range(0, len(titles), 128)]
    return titles, batches


@pytest.mark.parametrize("name", ["clean", "mixed"])
def test_reference_files_are_what_the_format_says(name):
    """The golden files themselves: format, map and titles (and, for the clean
    input, every sequence) as described."""
    check_db(os.path.join(GOLD, name), os.path.join(GOLD, name + ".fa"),
             exact_sequences=(name == "clean"))


def test_reference_mixed_batches_are_misplaced():
    """The reference's mixed-length batches pad at shifted positions
    (src/Bucket.cpp:123-130): the known defect this writer does not keep."""
    base = os.path.join(GOLD, "mixed")
    recs = {t: s for t, s in parse_fasta(base + ".fa")}
    titles = open(base + ".title").read().split("\n")[:-1]
    slots = [s for _, _, ss in seq_batches(base + ".seq") for s in ss]
    wrong = sum(not slots[i].startswith(padded_codes(recs[t])) for i, t in enumerate(titles))
    assert wrong > 100


@pytest.mark.parametrize("name", ["clean", "mixed"])
def test_writer_matches_reference(tmp_path, name):
    base = str(tmp_path / name)
    fa = os.path.join(GOLD, name + ".fa")
    records, slots = F.make(fa, base)
    assert records == len(parse_fasta(fa)) and slots % 128 == 0
    for ext in ("map", "title"):
        assert open(base + "." + ext, "rb").read() == open(os.path.join(GOLD, f"{name}.{ext}"), "rb").read()
    ours = seq_batches(base + ".seq")
    ref = seq_batches(os.path.join(GOLD, name + ".seq"))
    assert [(n, L) for n, L, _ in ours] == [(n, L) for n, L, _ in ref]
    if name == "clean":
        assert open(base + ".seq", "rb").read() == open(os.path.join(GOLD, "clean.seq"), "rb").read()
    else:
        for (n, L, a), (_, _, b) in zip(ours, ref):
            lens = {len(x.rstrip(bytes([23]))) for x in a}
            if n > 128 or len(lens) == 1:                        # single-length batches
                assert a == b
    check_db(base, fa)                                          # every sequence exact


@pytest.mark.parametrize("name", ["clean", "mixed"])
def test_reader_on_reference_files(name):
    base = os.path.join(GOLD, name)
    recs = {t: s for t, s in parse_fasta(base + ".fa")}
    titles = open(base + ".title").read().split("\n")[:-1]
    with F.DbFile(base) as f:
        assert f.records == len(titles) and f.slots % 128 == 0
        db = f.seqdb()
        for i, t in enumerate(titles):
            assert f.title(i) == t
            if name == "clean":
                want = np.frombuffer(padded_codes(recs[t]).rstrip(b"\x17"), np.uint8)
                assert np.array_equal(db.seq(i), want), t
        for i in range(len(titles), f.slots):
            assert f.title(i) == "" and db.lengths[i] == 0


def random_fasta(path, rng, n, crlf=False):
    lines = []
    for i in range(n):
        L = int(rng.choice([0, 1, 3, 4, 5, 17, 64, 128, 200, 333]) if i % 3 else rng.integers(0, 400))
        s = "".join(rng.choice(list("ACDEFGHIKLMNPQRSTVWYacdefXBZJ"), size=L))
        lines.append(f">r{i} x")
        w = int(rng.integers(10, 80))
        lines += [s[k:k + w] for k in range(0, L, w)]
        if i % 7 == 0:
            lines.append("")
    with open(path, "w", newline="") as f:
        f.write(("\r\n" if crlf else "\n").join(lines) + "\n")


@pytest.mark.parametrize("seed,n,flush,crlf", [(1, 700, 0, False), (2, 300, 500, True),
                                               (3, 129, 64, False), (4, 1, 0, False)])
def test_round_trip(tmp_path, seed, n, flush, crlf):
    rng = np.random.Generator(np.random.PCG64(seed))
    fa = str(tmp_path / "db.fa")
    random_fasta(fa, rng, n, crlf)
    base = str(tmp_path / "db")
    records, slots = F.make(fa, base, flush_bytes=flush)
    assert records == n
    titles, batches = check_db(base, fa)
    recs = dict(parse_fasta(fa))
    with F.DbFile(base) as f:
        db = f.seqdb()
        for i, t in enumerate(titles):
            assert f.title(i) == t
            real = np.frombuffer(padded_codes(recs[t]).rstrip(b"\x17"), np.uint8)
            assert np.array_equal(db.seq(i), real)


def test_fasta_reader(tmp_path):
    p = tmp_path / "q.fa"
    p.write_bytes(b"\n>q1 first\r\nACD EF\r\nghI\r\n\r\n>q2\nJ*@\n>q3 empty\n")
    recs = F.read_fasta(str(p))
    assert [t for t, _ in recs] == [">q1 first", ">q2", ">q3 empty"]
    assert list(recs[0][1]) == [ref_code(c) for c in "ACDEFGHI"]
    assert list(recs[1][1]) == [0xFF, 0xFF, 23]
    assert len(recs[2][1]) == 0
    bad = tmp_path / "bad.fa"
    bad.write_text("ACDE\n>x\nA\n")
    with pytest.raises(Exception, match="'>'"):
        F.read_fasta(str(bad))
    empty = tmp_path / "e.fa"
    empty.write_text("")
    assert F.read_fasta(str(empty)) == []


def test_dbmaker_tool(tmp_path):
    exe = os.path.join(ROOT, "swhybrid_amd", "dbmaker")
    out = subprocess.run([exe, os.path.join(GOLD, "clean.fa"), str(tmp_path / "c")],
                         capture_output=True, text=True, check=True)
    assert "total sequence num is" in out.stdout
    for ext in ("seq", "map", "title"):
        assert (tmp_path / f"c.{ext}").read_bytes() == open(os.path.join(GOLD, f"clean.{ext}"), "rb").read()
    bad = subprocess.run([exe], capture_output=True, text=True)
    assert bad.returncode == 2


def test_open_errors(tmp_path):
    with pytest.raises(Exception, match="cannot open"):
        F.DbFile(str(tmp_path / "missing"))
    (tmp_path / "t.seq").write_bytes(struct.pack("<QQ", 10**9, 0))
    with pytest.raises(Exception, match="end offset"):
        F.DbFile(str(tmp_path / "t"))
    (tmp_path / "u.seq").write_bytes(struct.pack("<QQ", 16 + 18 + 4, 0) + b"$" +
                                     struct.pack("<QQ", 1, 8) + b"#" + b"\x00" * 4)
    with pytest.raises(Exception, match="larger than the file"):
        F.DbFile(str(tmp_path / "u"))
