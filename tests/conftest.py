import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Builds the oracle when absent; brings libxbfs.so and ./bfs up to date
    with their sources (make's dependency check: a library older than any
    source is rebuilt, an up-to-date one is left alone)."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    from xbfs_amd import build as B
    B.build(force=False)
    yield
