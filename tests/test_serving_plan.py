"""The serving loop's budgeted-round planning on the host: every slice stays
inside its block kind (pair blocks [0, n_pairs), lane blocks [n_pairs,
n_blocks)), starts at the request's cursor and advances it, and each
request's slices cover every block exactly once, over 2,000 random block
tables and arrival patterns.  CPU only (tests/serving_plan_check.cpp against
swhybrid_amd/csrc/swh_plan.h)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_budgeted_plan_ranges(tmp_path):
    exe = tmp_path / "plan_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I",
                    os.path.join(ROOT, "swhybrid_amd", "csrc"),
                    os.path.join(ROOT, "tests", "serving_plan_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe), "2000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
