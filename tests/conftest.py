import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


class Sphere:
    """G sphere for a synthetic Si cell as the oracle computes it."""

    def __init__(self, ncell, ecut=20.0, ecutdens=80.0, cell=None):
        import oracle as O
        from qxmd_amd.system import diamond_si
        self.cell = diamond_si(ncell) if cell is None else cell
        self.nd = O.get_mesh(self.cell.hcell, ecutdens)
        s = O.strvec(self.cell.hcell, ecutdens, ecut)
        self.__dict__.update(s)
        self.ijk = O.ijkgd(s["nga"], s["ngb"], s["ngc"], s["nplw5"], self.nd)
        self.dkg = O.dkgnrm(s["recnrm"], s["nplwc"])[: 2 * (s["nplw"] + 1)]
        self.nplwex = 2 * (s["nplw"] + 1)
        self.N = int(np.prod(self.nd))


_SPHERES = {}


@pytest.fixture(scope="session")
def si8():
    if 1 not in _SPHERES:
        _SPHERES[1] = Sphere(1)
    return _SPHERES[1]


@pytest.fixture(scope="session")
def si64():
    if 2 not in _SPHERES:
        _SPHERES[2] = Sphere(2)
    return _SPHERES[2]
