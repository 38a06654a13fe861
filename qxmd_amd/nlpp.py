"""Tables of the nonlocal (Kleinman-Bylander) pseudopotential term that the
reference's setqlm reads from its nlpp_variables module (Sources/ppkb.f90:108-191,
allocated at 636-644 and 736-762), held as numpy arrays with the Fortran
dimensions reversed (C order), so that each array's memory is the reference's
column-major layout:

    lmx (ntype)                      number of projectors per type
    clalp (ntype, lmmax)             KB energies (Ry), vclalp = clalp / volume
    tbqlm, tbqlma (ntype, lnref, 551) radial table on |G| = m * dlqlm, m = 0..550,
                                     and its quadratic-interpolation term
    dlqlm (ntype)                    table step (bohr^-1)
    itola, itolm, itolmf (ntype, lmmax)  projector -> radial table row,
                                     angular row, real (1) / imaginary (2) angular part
    litoll (ntype, lmmax)            projector parity: True for even l
    qlmr, qlmi (nqlmmx, nplw + 1)    angular factors per G (Gamma point)

``model_tables`` builds s + p tables for silicon-like projectors, the same
functions the synthetic workloads use."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MXQLM = 550  # table length - 1 (Sources/ppkb.f90:140)


@dataclass
class NlppTables:
    lmx: np.ndarray
    clalp: np.ndarray
    tbqlm: np.ndarray
    tbqlma: np.ndarray
    dlqlm: np.ndarray
    itola: np.ndarray
    itolm: np.ndarray
    itolmf: np.ndarray
    litoll: np.ndarray
    qlmr: np.ndarray
    qlmi: np.ndarray

    @property
    def ntype(self):
        return len(self.lmx)

    @property
    def lmmax(self):
        return self.itola.shape[1]

    @property
    def lnref(self):
        return self.tbqlm.shape[1]


def interp_term(tb):
    """tbqlma for a radial table tb(m): 2 (tb(m) - 2 tb(m+1) + tb(m+2)), the
    second difference of the quadratic through m, m+1, m+2 that setqlm's
    interpolation uses (Sources/nlkbpp.f90:2207-2214)."""
    a = np.zeros_like(tb)
    a[..., :-2] = 2.0 * (tb[..., :-2] - 2.0 * tb[..., 1:-1] + tb[..., 2:])
    return a


def model_tables(gx, gy, gz, ntype=1, rc=1.2, dl=0.02, clalp=(-1.6, 0.9)):
    """s + p projector tables for G vectors (gx, gy, gz), G index 0..nplw:
    beta_0(g) = exp(-(g rc)^2 / 2), beta_1(g) = g rc exp(-(g rc)^2 / 2), real
    spherical harmonics Y_00 = 1 / sqrt(4 pi), Y_1m = sqrt(3 / 4 pi) (x, y, z) / |G|.
    Types differ by a scaled rc.  lmx = 4 (s, px, py, pz); lnref = 2."""
    gx, gy, gz = (np.asarray(v, np.float64) for v in (gx, gy, gz))
    g = np.sqrt(gx * gx + gy * gy + gz * gz)
    inv = np.where(g > 0, 1.0 / np.where(g > 0, g, 1.0), 0.0)
    nplw1 = len(g)
    qlmr = np.zeros((4, nplw1))
    qlmr[0] = 1.0 / np.sqrt(4 * np.pi)
    c1 = np.sqrt(3.0 / (4 * np.pi))
    qlmr[1], qlmr[2], qlmr[3] = c1 * gx * inv, c1 * gy * inv, c1 * gz * inv
    qlmi = np.zeros_like(qlmr)
    mgrid = np.arange(MXQLM + 1) * dl
    tb = np.zeros((ntype, 2, MXQLM + 1))
    for it in range(ntype):
        r = rc * (1.0 + 0.15 * it)
        e = np.exp(-0.5 * (mgrid * r) ** 2)
        tb[it, 0] = e
        tb[it, 1] = mgrid * r * e
    lm = np.tile(np.array([[1, 2, 3, 4]], np.int32), (ntype, 1))
    return NlppTables(
        lmx=np.full(ntype, 4, np.int32),
        clalp=np.tile(np.array([[clalp[0], clalp[1], clalp[1], clalp[1]]]), (ntype, 1)),
        tbqlm=tb, tbqlma=interp_term(tb), dlqlm=np.full(ntype, dl),
        itola=np.tile(np.array([[1, 2, 2, 2]], np.int32), (ntype, 1)),
        itolm=lm, itolmf=np.ones((ntype, 4), np.int32),
        litoll=np.tile(np.array([[1, 0, 0, 0]], np.int32), (ntype, 1)),
        qlmr=qlmr, qlmi=qlmi)
