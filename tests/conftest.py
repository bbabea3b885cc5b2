import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libswh.so on the GPU)")


def _flush_case(item):
    # name the case before a GPU step so a crashed run still says which one
    if item.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu-case] {item.nodeid}\n")
        sys.stdout.flush()


@pytest.hookimpl(tryfirst=True)
def pytest_runtest_setup(item):
    _flush_case(item)
