"""The threaded host code (DataPool and its fill threads, the CPU worker's aux
thread and scan team, RecalcWorker, ResultList merges, the packed-database
scan, the database files) under ThreadSanitizer and under AddressSanitizer +
UndefinedBehaviorSanitizer: tests/sanitize/host_stress.cpp, built by
`make -C swhybrid_amd/csrc sanitize SAN=...`, runs pipelines from two caller
threads at once and checks every score against the direct scan.  CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "swhybrid_amd", "csrc")


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_host_code_under_sanitizer(san, tmp_path):
    subprocess.run(["make", "-s", "-C", CSRC, "sanitize", f"SAN={san}"], check=True)
    exe = os.path.join(CSRC, "_obj", "san_" + san.replace(",", "_"), "host_stress")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1",
               ASAN_OPTIONS="detect_leaks=1 halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1 halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env,
                       timeout=600)
    report = [l for l in r.stderr.splitlines()
              if "Sanitizer" in l or "runtime error" in l]
    assert r.returncode == 0 and r.stdout.strip().endswith("ok") and not report, \
        (r.stdout[-2000:] + r.stderr[-4000:])
