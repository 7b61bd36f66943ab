import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


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
