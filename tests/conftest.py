import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "femeko.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def _print_case(item):
    # name the case before every GPU test so a lost box still tells which one ran
    if item.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu case] {item.nodeid}\n")
        sys.stdout.flush()


def pytest_runtest_setup(item):
    _print_case(item)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
