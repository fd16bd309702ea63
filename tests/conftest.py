import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")


@pytest.fixture(scope="session")
def engine():
    """Initialise the HIP engine on device 0. Fails (does not skip) when no GPU is usable."""
    import fomo_amd as F
    F.init(0)
    yield F


@pytest.fixture(scope="session")
def fe9_table(tmp_path_factory):
    import fomo_amd as F
    from oracle import oracle as O
    p = str(tmp_path_factory.mktemp("tables") / "fe9_171.dat")
    F.gen_goft_table(p, ion="fe_9", lambda0=171.0, atweight=56.0, nt=101, nrho=41, peak_logt=5.9, sigma=0.15)
    return p, O.read_table(p)


@pytest.fixture(scope="session")
def c0_data():
    """Config 0: structured 64^3 loop cube (seed 1)."""
    import fomo_amd as F
    return F.gen_loop_cube(64, seed=1)
