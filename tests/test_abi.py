"""The C ABI library loads and exports every symbol include/swh.h declares;
host-only entry points behave without a GPU.  CPU only."""
import ctypes as C
import os
import re

import numpy as np

import oracle as O
from swhybrid_amd import _native as N
from swhybrid_amd.engine import topk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(name):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(swh\w+)\s*\(", text))


def test_header_and_exports_agree():
    declared = header_functions("swh.h")
    assert declared == set(N.EXPORTS)
    lib = N.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_no_device_is_a_clean_error():
    n = C.c_int(-1)
    assert N.lib().swh_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        m = np.zeros(1024, dtype=np.int32)
        rc = N.lib().swh_ctx_create(0, N.ptr(m), 10, 2, C.byref(h))
        assert rc != 0 and not h.value
        assert b"device" in N.lib().swh_last_error()


def test_bad_arguments_are_rejected():
    h = C.c_void_p()
    bad = np.full(1024, 200, dtype=np.int32)
    assert N.lib().swh_ctx_create(0, N.ptr(bad), 10, 2, C.byref(h)) != 0
    assert N.lib().swh_ctx_create(0, None, 10, 2, C.byref(h)) != 0
    assert N.lib().swh_topk(None, 5, 3, None, None) != 0


def test_host_topk_matches_resultlist():
    rng = np.random.Generator(np.random.PCG64(1))
    cases = [(10000, 20, 60), (50, 10, 3), (7, 10, 5), (0, 4, 1), (100_000, 100, 70000)]
    cases += [(int(rng.integers(0, 3000)), int(rng.integers(0, 40)), int(rng.integers(1, 60)))
              for _ in range(100)]                       # block skipping across ties
    for n, k, hi in cases:
        s = rng.integers(-5, hi, size=n).astype(np.int32)
        a = topk(s, k)
        b = O.topk(s, k)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (n, k, hi)


def test_host_topk_large_vectors_in_parts():
    """Score vectors of a million and more are taken in parts on several
    threads; the result is the same total order (score desc, index asc)."""
    rng = np.random.Generator(np.random.PCG64(2))
    for n, hi, k in ((1_500_000, 90, 20), (2_000_003, 4, 50), (600_000, 70000, 10),
                     (524_289, 2, 100), (1_048_576, 1, 5)):
        s = rng.integers(-5, hi, size=n).astype(np.int32)
        a = topk(s, k)
        b = O.topk(s, k)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (n, hi, k)


def test_struct_mirrors_match_the_header(tmp_path):
    """Every ctypes mirror of a struct in include/swh.h and
    include/swhybrid_host.h has the C compiler's size and field offsets."""
    import subprocess
    mirrors = {"swh_stats": N.SwhStats, "swh_search_opts": N.SwhSearchOpts,
               "swh_db_opts": N.SwhDbOpts, "swh_tune_entry": N.SwhTuneEntry,
               "swh_tune_result": N.SwhTuneResult, "swh_server_opts": N.SwhServerOpts,
               "swh_request_info": N.SwhRequestInfo, "swh_server_stats": N.SwhServerStats,
               "swhy_pool_opts": N.SwhyPoolOpts, "swhy_pool_report": N.SwhyPoolReport}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "swh.h"',
             '#include "swhybrid_host.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if line:
            c, f, v = line.split()
            got[(c, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
