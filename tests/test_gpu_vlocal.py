"""GPU parity of vlocal_g (Sources/nlkbpp.f90:862-1032) through the C-ABI
against the oracle: the local pseudopotential from the device axis tables of
stscfn_pw into the SCF glue's external potential.  FP64, same summation order
per G as the reference; the transform's round-off bounds the grid values:
1e-12 of the largest value."""
import numpy as np
import pytest

import conftest
import oracle as O
from qxmd_amd import pw

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("llk", [(0, 0), (0, 1)])
def test_vlocal_g_device(si64, llk):
    sp = si64
    f = conftest.flocal_case(sp, seed=12)
    gs = pw.GSphere(sp.cell.hcell)
    ctx = pw.PlaneWave(gs, 4)
    rx = np.zeros(sp.nplw5 + 1)
    rx[1:] = 4 * np.pi / sp.recnrm[1: sp.nplw5 + 1]
    v_other = np.random.default_rng(2).standard_normal(sp.N) * 0.1
    ctx.scf_set_grid(rx, v_other)
    ctx.set_lopp(f["svco"])
    n5 = sp.nplw5
    ctx.stscfn_pw(f["frac"], sp.cell.hcell, sp.nd, sp.nd, sp.nga[: n5 + 1], sp.ngb[: n5 + 1],
                  sp.ngc[: n5 + 1], sp.nplw, f["kmax"], copy_out=False)
    ref, thr_ref = O.vlocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], list(llk),
                              f["natom"], sp.nga, sp.ngb, sp.ngc, f["svco"], sp.cell.volume,
                              f["kmax"], f["tables"], sp.ijk, sp.nd)
    args = (f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], list(llk), f["natom"], sp.cell.volume)
    v, thr = ctx.vlocal_g(*args, accumulate=True)
    assert _rel(thr, thr_ref) < 1e-13
    assert _rel(v - v_other, ref) < 1e-12
    v, _ = ctx.vlocal_g(*args)
    assert _rel(v, ref) < 1e-12


def test_vlocal_feeds_the_potential(si8):
    """After vlocal_g, scf_potential's glocal uses the new vext."""
    sp = si8
    f = conftest.flocal_case(sp, seed=3)
    gs = pw.GSphere(sp.cell.hcell)
    ctx = pw.PlaneWave(gs, 4)
    rx = np.zeros(sp.nplw5 + 1)
    rx[1:] = 4 * np.pi / sp.recnrm[1: sp.nplw5 + 1]
    ctx.scf_set_grid(rx)
    ctx.set_lopp(f["svco"])
    n5 = sp.nplw5
    ctx.stscfn_pw(f["frac"], sp.cell.hcell, sp.nd, sp.nd, sp.nga[: n5 + 1], sp.ngb[: n5 + 1],
                  sp.ngc[: n5 + 1], sp.nplw, f["kmax"], copy_out=False)
    rh = np.zeros(2 * (n5 + 1))
    rh[0] = 8.0 / sp.cell.volume
    rh[2:] = f["rhgr"][2:] * 0.01
    ctx.scf_put("rinr", rh)
    ctx.scf_put("rhgr", rh)
    v, _ = ctx.vlocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], [0, 0], f["natom"],
                        sp.cell.volume)
    ctx.mxchgg_alloc(imxchg=0)
    ctx.pwpmxchgg(0.8, 0.6)
    ctx.scf_potential()
    vh, vx = ctx.scf_get("vhar"), ctx.scf_get("vexc")
    assert _rel(ctx.scf_get("glocal"), vh + v + vx) < 1e-13
