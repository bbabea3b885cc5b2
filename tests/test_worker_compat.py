"""HIPWorker drops into the reference's own tree: its two files compile against
the reference's Worker.h (src/Worker.h:82-249) with the reference's compiler
flags (Makefiles/Makefile.cuda:19,38), and every Worker symbol the object
needs is defined by the reference's own Worker.cpp.  CPU only; skipped on a
machine without the reference checkout."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
HOST = os.path.join(ROOT, "swhybrid_amd", "csrc", "host")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "Worker.h")),
                                reason="reference checkout not present")

# the reference's CXXFLAGS for its AVX2 build, plus the C ABI's include dir
FLAGS = ["-O3", "-fopenmp", "-pthread", "-mavx2", "-DWITH_AVX2", "-I" + REF_SRC,
         "-I" + os.path.join(ROOT, "include")]


def in_reference_tree(tmp_path):
    """The maintainer's step: HIPWorker.{h,cpp} copied beside the reference's
    sources (here into a scratch directory searched before src/)."""
    for f in ("HIPWorker.h", "HIPWorker.cpp"):
        shutil.copy(os.path.join(HOST, f), tmp_path / f)
    return tmp_path / "HIPWorker.cpp"


def test_hipworker_compiles_against_reference_worker(tmp_path):
    src = in_reference_tree(tmp_path)
    r = subprocess.run(["g++", *FLAGS, "-fsyntax-only", "-I" + str(tmp_path), str(src)],
                       capture_output=True, text=True)
    errors = [l for l in r.stderr.splitlines() if "error" in l]
    assert r.returncode == 0 and not errors, r.stderr[-3000:]
    # the reference's Worker.h was the one included, not this repository's
    r = subprocess.run(["g++", *FLAGS, "-I" + str(tmp_path), "-M", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0
    deps = r.stdout.replace("\\\n", " ").split()
    assert os.path.join(REF_SRC, "Worker.h") in deps
    assert not any(d.startswith(HOST) for d in deps)


def undefined_and_defined(obj):
    out = subprocess.run(["nm", "-C", str(obj)], capture_output=True, text=True, check=True).stdout
    und, dfn = set(), set()
    for line in out.splitlines():
        m = re.match(r"\s*([0-9a-f]*)\s+([A-Za-z])\s+(.*)$", line)
        if not m:
            continue
        (und if m.group(2) == "U" else dfn).add(m.group(3))
    return und, dfn


def test_hipworker_links_against_reference_worker_cpp(tmp_path):
    src = in_reference_tree(tmp_path)
    hip_o, ref_o = tmp_path / "HIPWorker.o", tmp_path / "Worker.o"
    subprocess.run(["g++", *FLAGS, "-w", "-I" + str(tmp_path), "-c", str(src), "-o", str(hip_o)],
                   check=True)
    subprocess.run(["g++", *FLAGS, "-w", "-c", os.path.join(REF_SRC, "Worker.cpp"), "-o",
                    str(ref_o)], check=True)
    need, _ = undefined_and_defined(hip_o)
    _, have = undefined_and_defined(ref_o)
    from_worker = {s for s in need if "Worker" in s and "HIPWorker" not in s}
    assert from_worker, need
    missing = sorted(s for s in from_worker if s not in have)
    assert not missing, missing
    # beyond the base class it needs only the C ABI of libswh.so
    swh = {s for s in need if s.startswith("swh_")}
    assert swh == {"swh_ctx_create", "swh_ctx_destroy", "swh_db_create", "swh_db_destroy",
                   "swh_last_error", "swh_search"}
