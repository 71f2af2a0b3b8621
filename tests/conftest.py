import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("misa-md_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

A_FE = 2.855
RC_FACTOR = 5.6 / 2.855


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def setfl_fe(tmp_path_factory):
    import setfl_gen
    p = tmp_path_factory.mktemp("pot") / "fe.eam.alloy"
    setfl_gen.write_setfl(str(p), ("Fe",))
    return str(p)


@pytest.fixture(scope="session")
def setfl_fecuni(tmp_path_factory):
    import setfl_gen
    p = tmp_path_factory.mktemp("pot3") / "fecuni.eam.alloy"
    setfl_gen.write_setfl(str(p), ("Fe", "Cu", "Ni"))
    return str(p)
