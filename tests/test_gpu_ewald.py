"""GPU parity of the reciprocal-space Ewald forces and energies (frcrec,
Sources/bulk.f90:1880) through the C-ABI against the oracle.  The ion sums
S(G) are formed in the reference's order; forces and energies are summed over
G in another order: 1e-12 of the largest force, 1e-12 relative on energies."""
import numpy as np
import pytest

import oracle as O
from qxmd_amd import pw

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _args(e):
    return (e["ntype"], e["nhk1"], e["nhk2"], e["zv"], e["nel"], e["natom"], e["iatoit"],
            e["iatmpt"])


@pytest.mark.parametrize("ncell,numk2", [(1, 400), (2, 3000)])
def test_frcrec_host_tables(ncell, numk2, si8, si64):
    import conftest
    sp = si8 if ncell == 1 else si64
    e = conftest.ewald_case(sp, numk2=numk2)
    ctx = pw.PlaneWave(pw.GSphere(sp.cell.hcell), 4)
    ctx.set_ewald(e["ew"])
    f0 = np.full((e["natom"], 3), 0.5)
    a0 = np.full(e["natom"], -0.5)
    got = ctx.frcrec(*_args(e), e["tables"], f0, 1.25, a0)
    ref = O.frcrec(*_args(e), e["ew"], e["tables"], f0, 1.25, a0)
    assert _rel(got[0] - f0, ref[0] - f0) < 1e-12
    assert abs(got[1] - ref[1]) < 1e-12 * abs(ref[1])
    assert _rel(got[2], ref[2]) < 1e-12


def test_frcrec_device_tables(si64):
    """Per-atom axis tables from the device stscfn_pw (kmax of the density
    sphere covers the Ewald list)."""
    import conftest
    sp = si64
    e = conftest.ewald_case(sp, numk2=3000)
    ctx = pw.PlaneWave(pw.GSphere(sp.cell.hcell), 4)
    ctx.set_ewald(e["ew"])
    n5 = sp.nplw5
    kmax = [int(np.abs(np.asarray(a[: n5 + 1])).max()) for a in (sp.nga, sp.ngb, sp.ngc)]
    assert all(kmax[a] >= e["kcsma"][a] for a in range(3))
    ctx.stscfn_pw(e["frac"], sp.cell.hcell, sp.nd, sp.nd, sp.nga[: n5 + 1], sp.ngb[: n5 + 1],
                  sp.ngc[: n5 + 1], sp.nplw, kmax, copy_out=False)
    got = ctx.frcrec(*_args(e))
    ref = O.frcrec(*_args(e), e["ew"], e["tables"])
    assert _rel(got[0], ref[0]) < 1e-12
    assert abs(got[1] - ref[1]) < 1e-12 * abs(ref[1])
    assert _rel(got[2], ref[2]) < 1e-12
