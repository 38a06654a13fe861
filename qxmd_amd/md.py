"""Velocity-Verlet updates of the reference's MD loop at constant energy
(Sources/remd.f90: normalized_a 6789-6830, vkick 5408-5430, updtps 5755-5783,
updtps_pbc 5786-5806; rciprl, Sources/reclat.f90:1089-1128) through the
C-ABI.  Scaled coordinates: x (natom, 3) in cell units, v the displacement per
MD step, a = acon H^-1 f with acon = dt^2 / (2 m) per species; species are
1-based; hcell (3, 3) has the lattice vectors as columns.  One step:

    a = normalized_a(f, species, hcell, acon); v = vkick(v, a, ...)   # after forces
    v = vkick(v, a, ...); x = updtps_pbc(updtps(x, v, ...))           # next step
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import check, load


def _rows(x):
    return np.array(x, dtype=np.float64, order="C", copy=True).reshape(-1, 3)


def _ints(a):
    return np.ascontiguousarray(np.asarray(a).astype(np.int32))


def acon(mass, dt):
    """dt^2 / (2 m) per species (Sources/remd.f90:3568-3571)."""
    return 0.5 * float(dt) * float(dt) / np.asarray(mass, np.float64)


def rciprl(hcell):
    b = np.zeros(9)
    v = C.c_double(0.0)
    check(load().qx_rciprl(np.ascontiguousarray(np.asarray(hcell, np.float64).T).ravel(), b,
                           C.byref(v)), "qx_rciprl")
    return b.reshape(3, 3).T.copy(), v.value


def normalized_a(frc, species, hcell, acon_):
    a = _rows(frc)
    ac = np.ascontiguousarray(acon_, np.float64)
    check(load().qx_normalized_a(a.shape[0], len(ac), a.reshape(-1), _ints(species),
                                 np.ascontiguousarray(np.asarray(hcell, np.float64).T).ravel(),
                                 ac), "qx_normalized_a")
    return a


def vkick(v, a, species, lfixion):
    out = _rows(v)
    lf = _ints(np.asarray(lfixion, bool))
    check(load().qx_vkick(out.reshape(-1), _rows(a).reshape(-1), _ints(species), out.shape[0],
                          len(lf), lf), "qx_vkick")
    return out


def updtps(x, v, species, lfixion):
    out = _rows(x)
    lf = _ints(np.asarray(lfixion, bool))
    check(load().qx_updtps(out.reshape(-1), _rows(v).reshape(-1), _ints(species), out.shape[0],
                           len(lf), lf), "qx_updtps")
    return out


def updtps_pbc(x):
    out = _rows(x)
    check(load().qx_updtps_pbc(out.reshape(-1), out.shape[0]), "qx_updtps_pbc")
    return out
