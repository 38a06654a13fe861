"""The reference's restart files through the library: QM_eig (svwfns /
rdwfns, Sources/savedd.f90:268-1172, 2441-2810) and QM_ion (svacon / rdacon,
Sources/savedd.f90:40-265, 3058-3143), written byte-for-byte as the reference
writes them.  Band arrays are (nband, nplwex) in the reference's G order;
atom arrays (natom, 3); prevr (3, natom, 3); hcell with the lattice vectors
as columns."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import QxError, check, f64, i32, load


def svwfns(fname, gdcr, eig, occ, lsreal8=True):
    g = f64(gdcr)
    nb, n = g.shape
    check(load().qx_svwfns(str(fname).encode(), g, n, n // 2 - 1, nb, f64(eig), f64(occ),
                           int(bool(lsreal8))), "qx_svwfns")


def rdwfns(fname, nplw, nband):
    """Returns (gdcr (nb, nplwex), eig, occ, nband in the file), nb =
    min(nband, the file's)."""
    n = 2 * (int(nplw) + 1)
    g = np.zeros((nband, n))
    eig, occ = np.zeros(nband), np.zeros(nband)
    nf = C.c_int(0)
    check(load().qx_rdwfns(str(fname).encode(), g, n, int(nplw), int(nband), eig, occ,
                           C.byref(nf)), "qx_rdwfns")
    nb = min(nband, nf.value)
    return g[:nb], eig[:nb], occ[:nb], nf.value


def svacon(fname, nhk1, nhk2, nstepcg, nstepmd, lclust, ratm, ifmd, frc, prevr, hcell):
    check(load().qx_svacon(str(fname).encode(), len(nhk1), i32(nhk1), i32(nhk2), int(nstepcg),
                           int(nstepmd), int(bool(lclust)), f64(ratm), int(ifmd), f64(frc),
                           f64(prevr), f64(np.asarray(hcell).T)), "qx_svacon")


def rdacon(fname, nhk1, nhk2):
    """Returns dict(nstepcg, nstepmd, lclust, ratm, ifmd, frc, prevr, hcell)."""
    natom = int(nhk2[-1])
    ratm, frc = np.zeros((natom, 3)), np.zeros((natom, 3))
    prevr, h = np.zeros((3, natom, 3)), np.zeros((3, 3))
    a, b, lc, imd = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    check(load().qx_rdacon(str(fname).encode(), len(nhk1), i32(nhk1), i32(nhk2), C.byref(a),
                           C.byref(b), C.byref(lc), ratm, C.byref(imd), frc, prevr, h),
          "qx_rdacon")
    return dict(nstepcg=a.value, nstepmd=b.value, lclust=bool(lc.value), ratm=ratm,
                ifmd=imd.value, frc=frc, prevr=prevr, hcell=h.T.copy())


__all__ = ["svwfns", "rdwfns", "svacon", "rdacon", "QxError"]
