"""vlocal_g (Sources/nlkbpp.f90:862-1032) in the oracle: the local
pseudopotential on the grid from the form factors svco and the structure
factors.  The reference routine transforms through rfft3 (FFTW), so it is not
driven through the reference build; the oracle is pinned by the structure
factor written out with numpy's exp, the G = 0 term as the grid mean, and by
flocal_g (itself bit-equal to the reference build): the local forces must be
minus the gradient of the local energy E = rdelv sum_r vext(r) rho(r) that
this potential gives, which fixes the sign of the phase, the conjugation, the
1 / rvol scale and the +-G scatter."""
import numpy as np
import pytest

import conftest
import oracle as O


def _vloc(sp, f, frac, llk=(0, 0)):
    tables = O.axis_tables(frac, f["kmax"])
    return O.vlocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], list(llk), f["natom"],
                      sp.nga, sp.ngb, sp.ngc, f["svco"], sp.cell.volume, f["kmax"], tables,
                      sp.ijk, sp.nd)


@pytest.mark.parametrize("llk", [(0, 0), (1, 0)])
def test_vlocal_structure_factor(si8, llk):
    f = conftest.flocal_case(si8, seed=2)
    vext, thr = _vloc(si8, f, f["frac"], llk)
    n5 = si8.nplw5
    hkl = np.stack([np.asarray(a[: n5 + 1]) for a in (si8.nga, si8.ngb, si8.ngc)], axis=1)
    want = np.zeros(n5 + 1, complex)
    v0 = 0.0
    for it in range(2):
        if llk[it]:
            continue
        ions = np.asarray(f["iatmpt"])[f["nhk1"][it] - 1: f["nhk2"][it]] - 1
        s = np.exp(2j * np.pi * hkl @ f["frac"][ions].T).sum(axis=1)
        want += f["svco"][it] / si8.cell.volume * np.conj(s)
        v0 += len(ions) * f["svco"][it][0] / si8.cell.volume
    want[0] = 0.0
    got = thr[0::2] + 1j * thr[1::2]
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    assert vext.mean() == pytest.approx(v0, rel=1e-12, abs=1e-14)


def test_vlocal_forces_are_minus_gradient(si8):
    f = conftest.flocal_case(si8, seed=4)
    rho = O.chgg2r(f["rhgr"], si8.ijk, si8.nd)
    rdelv = si8.cell.volume / si8.N
    h = si8.cell.hcell

    def energy(frac):
        v, _ = _vloc(si8, f, frac)
        return rdelv * float(np.dot(v, rho))

    floc = O.flocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], [0, 0], f["natom"], si8.nga,
                      si8.ngb, si8.ngc, si8.gx, si8.gy, si8.gz, f["rhgr"], f["svco"], f["kmax"],
                      O.axis_tables(f["frac"], f["kmax"]))
    d = 1e-4
    for ia, ax in ((0, 0), (3, 1), (6, 2)):
        dt = np.zeros(3)
        dt[ax] = d
        ds = np.linalg.solve(h, dt)
        fp, fm = f["frac"].copy(), f["frac"].copy()
        fp[ia] += ds
        fm[ia] -= ds
        grad = (energy(fp) - energy(fm)) / (2 * d)
        assert -grad == pytest.approx(floc[ia, ax], rel=1e-6, abs=1e-9 * np.abs(floc).max())


def test_vlocal_of_gaussian_wells_is_the_model_potential(si8):
    """Form factors of Gaussian wells give back the model local potential on
    the grid (what the MD-step measurement relies on), up to the periodic
    images the model's minimum-image sum leaves out."""
    from qxmd_amd.system import model_local_potential
    sp = si8
    natom = sp.cell.tau.shape[0]
    depth, w, shift = 1.2, 1.1, 0.4
    g2 = np.asarray(sp.recnrm[: sp.nplw5 + 1])
    svco = -depth * (2 * np.pi * w * w) ** 1.5 * np.exp(-0.5 * g2 * w * w)
    svco[0] += shift * sp.cell.volume / natom
    frac = np.linalg.solve(sp.cell.hcell, sp.cell.tau.T).T
    kmax = [int(np.abs(np.asarray(a[: sp.nplw5 + 1])).max()) for a in (sp.nga, sp.ngb, sp.ngc)]
    v, _ = O.vlocal_g(1, [1], [natom], np.arange(1, natom + 1), [0], natom, sp.nga, sp.ngb,
                      sp.ngc, svco[None, :], sp.cell.volume, kmax, O.axis_tables(frac, kmax),
                      sp.ijk, sp.nd)
    assert np.abs(v - model_local_potential(sp, sp.cell)).max() < 1e-4
