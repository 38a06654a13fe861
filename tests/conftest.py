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


def nlpp_case(sp, ntype=2):
    """Inputs of stscfn_pw / setnlc_g for a Sphere: fractional atom positions
    (periodic cell, FD mesh = FFT mesh), kmax per axis over the nplw5 G list,
    s + p projector tables for ntype types splitting the atoms in equal runs."""
    from qxmd_amd import nlpp
    cell = sp.cell
    frac = np.linalg.solve(cell.hcell, cell.tau.T).T
    n5 = sp.nplw5 + 1
    g = [np.asarray(a[:n5], np.int32) for a in (sp.nga, sp.ngb, sp.ngc)]
    kmax = [int(np.abs(a).max()) for a in g]
    nd = np.array(sp.nd, np.int32)
    nw = sp.nplw + 1
    t = nlpp.model_tables(sp.gx[:nw], sp.gy[:nw], sp.gz[:nw], ntype=ntype)
    natom = len(frac)
    per = natom // ntype
    nhk1 = np.array([1 + per * it for it in range(ntype)], np.int32)
    nhk2 = np.array([per * (it + 1) if it < ntype - 1 else natom for it in range(ntype)], np.int32)
    return dict(frac=frac, nd=nd, g=g, kmax=kmax, tables=t, nhk1=nhk1, nhk2=nhk2, natom=natom,
                recnrm=np.asarray(sp.recnrm[:nw]), nplwcs=sp.nplw)


def flocal_case(sp, seed=3):
    """Inputs of flocal_g for a Sphere: two types splitting the atoms, iatmpt a
    fixed permutation, a random density on the density sphere and random
    local form factors svco (ntype, nplw5 + 1), axis tables of the atoms."""
    import oracle as O
    c = nlpp_case(sp)
    natom = c["natom"]
    n5 = sp.nplw5
    kmax = [int(np.abs(np.asarray(a[: n5 + 1])).max()) for a in (sp.nga, sp.ngb, sp.ngc)]
    rng = np.random.default_rng(seed)
    rhgr = rng.standard_normal(2 * (n5 + 1)) * 0.01
    rhgr[1] = 0.0
    svco = rng.standard_normal((2, n5 + 1))
    iatmpt = np.roll(np.arange(1, natom + 1), 3)
    return dict(ntype=2, nhk1=c["nhk1"], nhk2=c["nhk2"], iatmpt=iatmpt, natom=natom, kmax=kmax,
                tables=O.axis_tables(c["frac"], kmax), rhgr=rhgr, svco=svco, frac=c["frac"])


def ewald_case(sp, numk2=400, gamma=0.35, seed=4):
    """Inputs of frcrec for a Sphere: the Ewald G list = the first numk2
    half-sphere G of strvec in units without 2 pi (as clmrcp builds them),
    recnrmex = 0 (no double grid), two types (zv 4 and 3) over a permuted
    iatmpt, per-ion axis tables cos/sin(2 pi k s) up to the list's kcsma."""
    import oracle as O
    c = nlpp_case(sp)
    natom = c["natom"]
    tp = 2 * np.pi
    m = slice(1, numk2 + 1)
    ew = dict(nga=np.asarray(sp.nga[m], np.int32), ngb=np.asarray(sp.ngb[m], np.int32),
              ngc=np.asarray(sp.ngc[m], np.int32), gx=np.asarray(sp.gx[m]) / tp,
              gy=np.asarray(sp.gy[m]) / tp, gz=np.asarray(sp.gz[m]) / tp,
              recnrm=np.asarray(sp.recnrm[m]) / tp ** 2, recnrmex=np.zeros(numk2 + 1),
              gamma=gamma, volume=sp.cell.volume)
    kcsma = [int(np.abs(ew[k]).max()) for k in ("nga", "ngb", "ngc")]
    rng = np.random.default_rng(seed)
    iatmpt = rng.permutation(natom) + 1
    iatoit = np.where(np.arange(natom) < natom // 2, 1, 2).astype(np.int32)
    zv = np.array([4.0, 3.0])
    nhk1 = np.array([1, natom // 2 + 1], np.int32)
    nhk2 = np.array([natom // 2, natom], np.int32)
    nel = int(sum((nhk2[i] - nhk1[i] + 1) * zv[i] for i in range(2)))
    frac = c["frac"]
    tables = O.axis_tables(frac[iatmpt - 1], kcsma)
    return dict(ntype=2, nhk1=nhk1, nhk2=nhk2, zv=zv, nel=nel, natom=natom, iatoit=iatoit,
                iatmpt=iatmpt, ew=ew, tables=tables, kcsma=kcsma, frac=frac)


def roteig_case(sp, nb=8, nprv=3, seed=1):
    """Current orthonormal bands and nprv previous sets near them."""
    import oracle as O
    from qxmd_amd.system import initial_wavefunctions
    cur, _ = O.schmidt(initial_wavefunctions(sp, nb, 25.0))
    rng = np.random.default_rng(seed)
    prev = np.stack([O.schmidt(cur + 0.1 * rng.standard_normal(cur.shape))[0]
                     for _ in range(nprv)])
    return cur, prev


def restart_case(sp, nb=5):
    """Bands, eigenvalues and occupations for QM_eig, and an atom
    configuration for QM_ion (two types over 8 atoms, a sheared cell)."""
    from qxmd_amd.system import initial_wavefunctions
    g = initial_wavefunctions(sp, nb, 25.0)
    eig = np.linspace(-0.5, 0.3, nb)
    occ = np.array([2.0, 2.0, 2.0, 1.5, 0.0][:nb])
    rng = np.random.default_rng(0)
    natom = 8
    h = np.diag([10.26, 11.0, 12.5])
    h[0, 1] = 0.3
    ion = dict(nhk1=[1, 5], nhk2=[4, 8], nstepcg=3, nstepmd=17, lclust=False,
               ratm=rng.uniform(size=(natom, 3)), ifmd=1, frc=rng.standard_normal((natom, 3)),
               prevr=rng.standard_normal((3, natom, 3)), hcell=h)
    return g, eig, occ, ion


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


def md_case(seed=11):
    """Inputs of the velocity-Verlet updates: 23 atoms of 3 species (the
    second fixed), a skewed cell, scaled positions partly outside [0, 1)
    (with the edge values 0, -0, +-1, +-2 and -2.5), displacements and
    Cartesian forces."""
    rng = np.random.default_rng(seed)
    n = 23
    species = np.array([1] * 8 + [2] * 5 + [3] * 10, np.int32)
    lfixion = np.array([False, True, False])
    hcell = np.array([[10.26, 0.4, -0.3], [0.2, 11.0, 0.5], [-0.1, 0.3, 9.7]])
    x = rng.uniform(-0.2, 1.2, (n, 3))
    x.reshape(-1)[:7] = [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, -2.5]
    v = rng.normal(0.0, 1e-3, (n, 3))
    f = rng.normal(0.0, 0.05, (n, 3))
    acon = np.array([0.5 * 40.0 ** 2 / 51196.0, 0.5 * 40.0 ** 2 / 21874.0,
                     0.5 * 40.0 ** 2 / 1837.0])
    return dict(species=species, lfixion=lfixion, hcell=hcell, x=x, v=v, f=f, acon=acon)
