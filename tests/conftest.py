import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "milc-qphix_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def qlib():
    """The HIP library, loaded once per session (GPU tests only)."""
    from pyqphix import lib as L
    return L.load()


_CURRENT = {}


def lattice_session(dims):
    """The library holds one global lattice (as the reference does); switch
    lattices by finalize + init."""
    from pyqphix.api import Session
    cur = _CURRENT.get("s")
    if cur is not None and cur.dims == tuple(dims):
        return cur
    if cur is not None:
        cur.close()
    print(f"[qphix] init lattice {dims}", flush=True)
    s = Session(dims)
    _CURRENT["s"] = s
    return s
