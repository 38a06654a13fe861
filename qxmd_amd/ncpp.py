"""Norm-conserving pseudopotential tables for a plane-wave sphere from the
radial tables the reference's set-up produces.

QXMD turns a pseudopotential into two kinds of data: radial tables that depend
on the cell only through the sphere's largest |G| (the KB table tbqlm of
nlkbpp_g, Sources/ppkb.f90:2494-3295; the local table rhocr of lotmpp_g,
Sources/ppkb.f90:2148-2476) and per-G arrays built from them and from the G
vectors.  This module rebuilds the per-G arrays:

* ``angular_factors``: qlmr / qlmi, the real spherical-harmonic factors
  nlkbpp_g stores per G for the Gamma point (the lyreal branch,
  Sources/ppkb.f90:2605-2660): rows l(l+1)/2 + m + 1 for m = 0..l, the m = 0
  row s_l Y_l0, the m > 0 rows s_l sqrt(2) (Re, Im) Y_lm, with s_l = +1 when
  l/2 is even and -1 otherwise, and Y_lm = N_lm P_l^m(cos t) e^{i m phi}
  without the Condon-Shortley sign, as the reference's sphrb2 tabulates them
  (Sources/ppvd.f90:193-330);
* ``form_factors``: svco = 4 pi v(|G|) - 2 zv recnrmex(G), with v read from
  lotmpp_g's table by the same three-point rule and recnrmex = 4 pi / |G|^2
  (set_recnrmex, Sources/ppkb.f90:2075-2128).

``nlpp_tables`` assembles the NlppTables that qx_set_nlpp_tables takes.
"""
from __future__ import annotations

import numpy as np

from .nlpp import NlppTables

MXQAB = 1000  # length - 1 of lotmpp_g's radial table (Sources/ppkb.f90:2139)


def _ylm(l, m, c1, s1, c2m, s2m):
    """N_lm P_l^m(cos t) (cos m phi, sin m phi) for l <= 3, m >= 0; c1, s1 =
    cos t, sin t; c2m, s2m = cos m phi, sin m phi."""
    pi = np.pi
    if l == 0:
        p = np.full_like(c1, np.sqrt(1.0 / (4.0 * pi)))
    elif l == 1:
        p = np.sqrt(3.0 / (4.0 * pi)) * c1 if m == 0 else np.sqrt(3.0 / (8.0 * pi)) * s1
    elif l == 2:
        p = {0: np.sqrt(5.0 / (16.0 * pi)) * (3.0 * c1 * c1 - 1.0),
             1: np.sqrt(15.0 / (8.0 * pi)) * s1 * c1,
             2: np.sqrt(15.0 / (32.0 * pi)) * s1 * s1}[m]
    elif l == 3:
        p = {0: np.sqrt(7.0 / (16.0 * pi)) * c1 * (5.0 * c1 * c1 - 3.0),
             1: np.sqrt(21.0 / (64.0 * pi)) * s1 * (5.0 * c1 * c1 - 1.0),
             2: np.sqrt(105.0 / (32.0 * pi)) * s1 * s1 * c1,
             3: np.sqrt(35.0 / (64.0 * pi)) * s1 * s1 * s1}[m]
    else:
        raise ValueError("l <= 3 only")
    return p * c2m, p * s2m


def angular_factors(gx, gy, gz, lmax):
    """(qlmr, qlmi), each (nqlmmx, nplw + 1) with nqlmmx = (lmax+1)(lmax+2)/2."""
    gx, gy, gz = (np.asarray(v, np.float64) for v in (gx, gy, gz))
    small = 1.0e-5
    gxy = np.sqrt(gx * gx + gy * gy)
    g = np.sqrt(gx * gx + gy * gy + gz * gz)
    big = g > small
    gs = np.where(big, g, 1.0)
    c1 = np.where(big, gz / gs, 1.0)
    s1 = np.where(big, gxy / gs, 0.0)
    inplane = big & (gxy > small)
    gxys = np.where(inplane, gxy, 1.0)
    c2 = np.where(inplane, gx / gxys, 1.0)
    s2 = np.where(inplane, gy / gxys, 0.0)
    nq = (lmax + 1) * (lmax + 2) // 2
    qr = np.zeros((nq, len(g)))
    qi = np.zeros((nq, len(g)))
    row = 0
    for l in range(lmax + 1):
        sgn = 1.0 if (l // 2) % 2 == 0 else -1.0
        cm, sm = np.ones_like(c2), np.zeros_like(s2)  # cos / sin of m phi
        for m in range(l + 1):
            yr, yi = _ylm(l, m, c1, s1, cm, sm)
            f = sgn if m == 0 else sgn * np.sqrt(2.0)
            qr[row] = f * yr
            qi[row] = f * yi if m > 0 else 0.0
            row += 1
            cm, sm = cm * c2 - sm * s2, sm * c2 + cm * s2
    return qr, qi


def recnrmex(recnrm5):
    """4 pi / |G|^2 over the density sphere, 0 at G = 0 (set_recnrmex)."""
    r = np.asarray(recnrm5, np.float64)
    out = np.zeros_like(r)
    out[1:] = 4.0 * np.arccos(-1.0) / r[1:]
    return out


def form_factors(recnrm5, rhocr, rhocra, zv):
    """svco(0..nplw5) of one species from lotmpp_g's radial table rhocr and its
    second-difference term rhocra; the table step follows from the sphere's
    largest |G| (recnrm5 is sorted, its last entry is the largest)."""
    r5 = np.asarray(recnrm5, np.float64)
    rhocr = np.asarray(rhocr, np.float64)
    rhocra = np.asarray(rhocra, np.float64)
    g1maxq = np.sqrt(r5[-1])
    dlqab = g1maxq * (1.0 + 10.0 / MXQAB) / MXQAB
    ak1 = np.sqrt(r5)
    m = 2 * np.trunc(ak1 / (2.0 * dlqab)).astype(np.int64)
    d = 0.5 * (ak1 / dlqab - m)
    rhps = d * ((d - 1.0) * rhocra[m] + rhocr[m + 2] - rhocr[m]) + rhocr[m]
    return 4.0 * np.pi * rhps - 2.0 * zv * recnrmex(r5)


def nlpp_tables(gx, gy, gz, species):
    """NlppTables for the sphere (gx, gy, gz over G = 0..nplw) and a list of
    species, each a dict with the radial tables of one type: lmx, clalp,
    itola, itolm, itolmf, litoll (length lmmax), tbqlm, tbqlma (lnref, 551),
    dlqlm, lmax (angular rows up to it)."""
    lmax = max(int(s["lmax"]) for s in species)
    qr, qi = angular_factors(gx, gy, gz, lmax)

    def stack(k, dt=None):
        return np.stack([np.asarray(s[k], dt) for s in species])

    return NlppTables(lmx=np.array([int(s["lmx"]) for s in species], np.int32),
                      clalp=stack("clalp", np.float64), tbqlm=stack("tbqlm", np.float64),
                      tbqlma=stack("tbqlma", np.float64),
                      dlqlm=np.array([float(s["dlqlm"]) for s in species]),
                      itola=stack("itola", np.int32), itolm=stack("itolm", np.int32),
                      itolmf=stack("itolmf", np.int32), litoll=stack("litoll", np.int32),
                      qlmr=qr, qlmi=qi)


def si_lda(fixture, tag):
    """The silicon species of the fixture tests/golden/si_ncpp_lda.npz (an
    np.load of it) for the cell tagged `tag` (tiny, si8, si64, si216, si512):
    a dict for nlpp_tables plus zv, rhocr, rhocra for form_factors."""
    f = fixture
    return dict(lmx=int(f["lmx"]), lmax=int(f["lmax"]), zv=float(f["zv"]),
                clalp=f["clalp"], itola=f["itola"], itolm=f["itolm"], itolmf=f["itolmf"],
                litoll=f["litoll"], tbqlm=f[f"{tag}_tbqlm"], tbqlma=f[f"{tag}_tbqlma"],
                dlqlm=float(f[f"{tag}_dlqlm"]), rhocr=f[f"{tag}_rhocr"],
                rhocra=f[f"{tag}_rhocra"])


def silicon_cell_inputs(gs, fixture, tag):
    """Everything the SCF needs from the silicon pseudopotential for one
    plane-wave sphere `gs` (qxmd_amd.pw.GSphere): the NlppTables of
    qx_set_nlpp_tables, the local form factors svco (1, nplw5 + 1) and the
    valence charges zv (1,)."""
    sp = si_lda(fixture, tag)
    nw = gs.nplw + 1
    t = nlpp_tables(gs.gx[:nw], gs.gy[:nw], gs.gz[:nw], [sp])
    svco = form_factors(gs.recnrm[: gs.nplw5 + 1], sp["rhocr"], sp["rhocra"], sp["zv"])
    return t, svco[None, :], np.array([sp["zv"]])
