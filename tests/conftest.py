import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pytest_collection_finish(session):
    """On a machine with a GPU, start torch's HIP runtime before any test
    creates an engine context: torch ships its own HIP runtime next to the
    system one that libinca_hip.so uses, and tests that generate records
    with torch after an engine exists found no device in torch when torch
    came second."""
    if not os.path.exists("/dev/kfd"):
        return
    if not any(item.get_closest_marker("gpu") for item in session.items):
        return
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


@pytest.fixture(scope="session")
def ch4():
    """The reference's CH4 inputs (tests/data of markelgtx/INCA_parallel)."""
    from inca_inputs import read_bas, read_dm2p, read_fchk
    g = os.path.join(GOLDEN, "ch4")
    mol = read_bas(os.path.join(g, "ch4_ops.bas"), read_fchk(os.path.join(g, "ch4.fchk")))
    ijkl, d = read_dm2p(os.path.join(g, "ch4.dm2p"))
    mol["ijkl"], mol["d"] = ijkl, d
    mol["dm2p_path"] = os.path.join(g, "ch4.dm2p")
    return mol


@pytest.fixture(scope="session")
def engine():
    from inca_parallel_amd import capi
    e = capi.Engine(0)
    yield e
    e.close()
