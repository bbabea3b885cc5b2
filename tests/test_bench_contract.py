"""The measurement script's contract that needs no device: one GPU only, and
the committed rocprofv3 counters it reports with the roofline."""
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""   # no JSON line for an unsupported run


def test_committed_counters_are_consistent():
    """The traffic and VALU figures bench.py reports come from
    profiles/k_scan_traffic.json: its means are the means of its per-query
    figures, and the counter summary it cites states the same mean."""
    sys.path.insert(0, ROOT)
    import bench

    traffic, ipp = bench.measured_counters()
    with open(os.path.join(ROOT, "profiles", "k_scan_traffic.json")) as f:
        t = json.load(f)
    per = t["per_launch_bytes"]
    assert sorted(per) == ["1000", "144", "2005", "464", "5478"]
    assert abs(traffic - sum(per.values()) / len(per)) <= 0.001 * traffic
    vals = list(t["valu_insts_per_cell_pair"].values())
    assert abs(ipp - sum(vals) / len(vals)) < 0.006
    # the recurrence needs 7.5 instructions per cell pair; the count cannot be below it
    assert bench.MIN_INSTS_PER_CELL_PAIR <= ipp < 8.5
    with open(os.path.join(ROOT, "profiles", "k_scan_plain_5M_counters.md")) as f:
        md = f.read()
    m = re.search(r"roofline\.traffic`\): ([\d,.]+) GB", md)
    assert m and abs(float(m.group(1).replace(",", "")) * 1e9 - traffic) <= 0.001 * traffic
    m = re.search(r"Mean over the five: ([\d.]+) instructions per cell pair", md)
    assert m and abs(float(m.group(1)) - ipp) < 0.006
