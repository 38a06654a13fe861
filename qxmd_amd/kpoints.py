"""Bloch bands at several k-points: the KB projector coefficients at k+G and
the Gamma-folding map between a cell's k-points and its supercell.

The reference's band algebra is Gamma-only (Sources/kpoints.f90:33-39); the
k-point H|psi> of the library (qx_kpoints_set, qx_khpsi_slot) takes complex
bands c_nk(G) on each k-point's own sphere |k+G|^2 <= E_cut.  This module
builds what that needs on the host:

* ``kb_projectors``: beta_p(k+G) for every KB projector p = (atom, l row),
  in the projector order of setnlc_g (atom-major, rows of the species in
  order, KB species only; Sources/nlkbpp.f90:1236-1279).  Each coefficient is
  setbgq / setqlm's Gamma expression (Sources/nlkbpp.f90:2171-2183,
  2207-2244) evaluated at the vector k+G: the radial table tbqlm read at
  |k+G| by the same three-point rule, the real spherical-harmonic factor of
  the direction of k+G (qxmd_amd.ncpp.angular_factors), the phase
  exp(-i (k+G).tau) and the factor -i on the rows whose litoll is 0 (the rows
  the Gamma table stores as (-sin, -cos)).  D_p = clalp / volume (vclalp).
* ``supercell_map``: for a supercell of m x m x m cells, which k-point of
  {0, 1/m, ..}^3 and which sphere point each supercell G' is (G' = m (k+G)),
  the folding under which the supercell's Gamma H|psi> equals the cell's
  H_k on every k-point.
"""
from __future__ import annotations

import numpy as np

from .ncpp import angular_factors


def k_cartesian(bvec, kfrac, hkl):
    """Cartesian k+G (npw, 3) for integer hkl (npw, 3); bvec has the
    reciprocal vectors (2 pi included) as columns."""
    h = np.asarray(hkl, np.float64).reshape(-1, 3)
    return (h + np.asarray(kfrac, np.float64)[None, :]) @ np.asarray(bvec, np.float64).T


def kb_projectors(tables, tau, species, bvec, kfrac, hkl, volume):
    """beta (npj, 2 npw) complex interleaved and D (npj,) for one k-point.

    tables: NlppTables (radial tables, index rows; its qlmr / qlmi are not
    used: the angular factors are evaluated at k+G here); tau (natom, 3)
    Cartesian positions (bohr); species (natom,) 0-based type of each atom;
    bvec reciprocal vectors (columns, 2 pi included); kfrac the k-point in
    units of bvec; hkl (npw, 3) the sphere; volume the cell volume."""
    kg = k_cartesian(bvec, kfrac, hkl)
    g = np.sqrt(np.sum(kg * kg, axis=1))
    lmax = 0
    while (lmax + 1) * (lmax + 2) // 2 < int(np.max(tables.itolm)):
        lmax += 1
    qr, qi = angular_factors(kg[:, 0], kg[:, 1], kg[:, 2], lmax)
    ntype, lmmax = tables.itola.shape
    betas, D = [], []
    tau = np.asarray(tau, np.float64)
    for ia in range(tau.shape[0]):
        it = int(species[ia])
        ph = np.exp(-1j * (kg @ tau[ia]))
        dl = float(tables.dlqlm[it])
        m = 2 * np.trunc(g / (2.0 * dl)).astype(np.int64)
        d = 0.5 * (g / dl - m)
        for l in range(int(tables.lmx[it])):
            tb = tables.tbqlm[it, int(tables.itola[it, l]) - 1]
            ta = tables.tbqlma[it, int(tables.itola[it, l]) - 1]
            zz = d * ((d - 1.0) * ta[m] + tb[m + 2] - tb[m]) + tb[m]
            q = qr if int(tables.itolmf[it, l]) == 1 else qi
            z = zz * q[int(tables.itolm[it, l]) - 1]
            b = z * ph if int(tables.litoll[it, l]) else -1j * z * ph
            betas.append(b)
            D.append(float(tables.clalp[it, l]) / volume)
    beta = np.zeros((len(betas), 2 * len(g)))
    for p, b in enumerate(betas):
        beta[p, 0::2] = b.real
        beta[p, 1::2] = b.imag
    return beta, np.array(D)


def gamma_folding_kpoints(m):
    """The m^3 k-points {0, 1/m, .., (m-1)/m}^3 (units of the cell's
    reciprocal vectors) that fold onto the Gamma point of the m x m x m
    supercell, k1 slowest."""
    r = np.arange(m) / m
    return np.array([[a, b, c] for a in r for b in r for c in r])


def supercell_map(sup_g, m, sphere_index):
    """For supercell G' (n, 3 integers, units of the supercell's reciprocal
    vectors = cell's / m): the k-point index (in gamma_folding_kpoints(m)
    order) and the sphere point of that k-point, G' = m (k + G).
    sphere_index: list over k-points of dicts (h, k, l) -> point.  Points
    absent from the k-point's sphere get -1."""
    sup_g = np.asarray(sup_g, np.int64)
    kint = np.mod(sup_g, m)
    kidx = (kint[:, 0] * m + kint[:, 1]) * m + kint[:, 2]
    g = (sup_g - kint) // m
    pt = np.array([sphere_index[int(ki)].get((int(a), int(b), int(c)), -1)
                   for ki, (a, b, c) in zip(kidx, g)], np.int64)
    return kidx, pt
