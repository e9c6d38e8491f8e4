import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "slow: long parity runs")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    """The CPU oracle is compiled once per session (test infrastructure)."""
    import oracle
    if not os.path.exists(oracle.LIB):
        oracle.build()
