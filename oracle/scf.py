"""CPU SCF of one cell with the oracle's routines: the outer loop of
pwlda_md (Sources/pwlda.f90:1628-3700) in the order the reference runs it.

TEST INFRASTRUCTURE ONLY (the checker of qxmd_amd.scf.Scf).  Per iteration:
the band step (qxo_scf_band_step: energy pass, subspace rotation, BKL CG with
the reference's defaults, schmidt, softchg), chgr2g of the output density,
Pulay mixing (mxchgg_pulay), and the potential V = v_H + v_ext + v_xc from the
mixed density (chgg2r, harting, xclda; pwpunifylc adds the three,
Sources/engrad.f90:667-799).  The energy tested for convergence is the
Harris-Foulkes energy of the band step's eigenvalues and the mixed density,
the same expression the device driver uses.  Every piece is pinned on its
own (see oracle.PINNING); this module only composes them.
"""
from __future__ import annotations

import numpy as np

import oracle as O


class OracleScf:
    def __init__(self, sph, natom, tables, svco, zv, nhk1=None, nhk2=None, itratn=10,
                 aslh=0.8, bslh=0.6, cmetric=20.0):
        """sph: a plane-wave sphere with cell, nd, ijk, dkg, recnrm, gx/gy/gz,
        nga/ngb/ngc, nplw, nplw5, nplwc (tests/conftest.py Sphere)."""
        self.s = sph
        self.natom = int(natom)
        self.t = tables
        self.svco = np.atleast_2d(np.asarray(svco, np.float64))
        self.ntype = self.svco.shape[0]
        self.zv = np.atleast_1d(np.asarray(zv, np.float64))
        self.nhk1 = np.array([1] if nhk1 is None else nhk1, np.int32)
        self.nhk2 = np.array([self.natom] if nhk2 is None else nhk2, np.int32)
        self.aslh, self.bslh = float(aslh), float(bslh)
        s = sph
        n5 = s.nplw5 + 1
        self.g = [np.asarray(a[:n5], np.int32) for a in (s.nga, s.ngb, s.ngc)]
        self.kmax = [int(np.abs(a).max()) for a in self.g]
        self.rvol = s.cell.volume
        self.rdelv = self.rvol / s.N
        self.rx = np.zeros(n5)
        self.rx[1:] = 4.0 * np.arccos(-1.0) / np.asarray(s.recnrm[1:n5])
        ncx = 2 * (s.nplwc + 1)
        q1 = O.factq1(cmetric, s.recnrm[1], s.recnrm[s.nplwc])
        self.mixer = O.PulayMixer(ncx, itratn, O.dkgnrm(s.recnrm, s.nplwc)[:ncx], self.rvol, q1)
        self.types = np.zeros(self.natom, np.int64)
        for it in range(self.ntype):
            self.types[self.nhk1[it] - 1: self.nhk2[it]] = it

    def set_positions(self, x):
        s = self.s
        nd = np.array(s.nd, np.int32)
        self.x = np.asarray(x, np.float64)
        yc, ys = O.stscfn_pw(self.x, s.cell.hcell, nd, nd, *self.g, s.nplw, self.kmax)
        self.eq, vcl = O.setnlc_g(self.t, self.nhk1, self.nhk2, self.natom, yc, ys,
                                  np.asarray(s.recnrm[: s.nplw + 1]), self.rvol)
        self.lmx_atom = np.asarray(self.t.lmx, np.int32)[self.types]
        self.vcl_atom = vcl[self.types]
        tables = O.axis_tables(self.x, self.kmax)
        self.tables = tables
        self.vext, _ = O.vlocal_g(self.ntype, self.nhk1, self.nhk2, np.arange(1, self.natom + 1),
                                  np.zeros(self.ntype, np.int32), self.natom, *self.g, self.svco,
                                  self.rvol, self.kmax, tables, s.ijk, s.nd)

    def _potential(self):
        s = self.s
        rho = O.chgg2r(self.rh, s.ijk, s.nd)
        vh = O.harting(self.rh, self.rx, s.ijk, s.nd)
        vx, _, e4 = O.xclda(rho, None, self.rdelv)
        self.V = vh + self.vext + vx
        self.energies = dict(ehar=0.5 * float(np.sum(vh * rho)) * self.rdelv,
                             elcl=float(np.sum(self.vext * rho)) * self.rdelv,
                             eexc=e4[0], evexc=e4[1], ecor=e4[2], evcor=e4[3])

    def start(self, psi, occ):
        s = self.s
        self.psi = np.asarray(psi, np.float64).copy()
        rho = O.softchg(self.psi, s.nplw, occ, self.rvol, s.ijk, s.nd)
        rh = O.chgr2g(rho, s.nplw5, s.ijk, s.nd)
        self.rh, self.rin = self.mixer.mix(rh, rh.copy(), self.aslh, self.bslh)
        self._potential()

    def hamiltonian(self):
        s = self.s
        return O.Hamiltonian(s.nplw, s.nd, s.ijk, self.V, s.dkg, self.eq, self.lmx_atom,
                             self.vcl_atom)

    harris_foulkes = staticmethod(lambda occ, eig, en: (
        float(np.sum(np.asarray(occ) * np.sort(eig))) - en["ehar"] + en["eexc"] + en["ecor"]
        - en["evexc"] - en["evcor"]))

    def iterate(self, occ, tol_ha_atom=1e-6, max_scf=60, reset_history=True):
        s = self.s
        hist = []
        for it in range(max_scf):
            r = self.hamiltonian().scf_band_step(self.psi, occ, self.rvol)
            if not r["ok"]:
                raise RuntimeError("schmidt: overlap not positive definite")
            self.psi = r["psi"]
            rh = O.chgr2g(r["rho"], s.nplw5, s.ijk, s.nd)
            self.rh, self.rin = self.mixer.mix(rh, self.rin, self.aslh, self.bslh,
                                               lreset=(reset_history and it == 0))
            self._potential()
            hist.append(self.harris_foulkes(occ, r["eig"], self.energies))
            if len(hist) > 1 and abs(hist[-1] - hist[-2]) / 2.0 / self.natom < tol_ha_atom:
                return it + 1, True, hist
        return max_scf, False, hist

    def finish(self, occ):
        H = self.hamiltonian()
        hp = H.hpsi(self.psi)
        hm, _ = O.hmatrix(self.psi, hp)
        w, u = O.eigh(hm)
        self.psi = O.untrycore(self.psi, u)
        return w, self.harris_foulkes(occ, w, self.energies)

    def forces(self, occ):
        s = self.s
        nw = s.nplw + 1
        _, sl = O.hckbpp_g(self.eq, self.lmx_atom, self.vcl_atom, self.psi,
                           np.zeros_like(self.psi))
        f = O.kbnlpp_frc2(self.eq, self.lmx_atom, self.vcl_atom, self.psi, occ, sl, s.gx[:nw],
                          s.gy[:nw], s.gz[:nw])
        n5 = s.nplw5 + 1
        f = f + O.flocal_g(self.ntype, self.nhk1, self.nhk2, np.arange(1, self.natom + 1),
                           np.zeros(self.ntype, np.int32), self.natom, *self.g,
                           s.gx[:n5], s.gy[:n5], s.gz[:n5], self.rh, self.svco, self.kmax,
                           self.tables)
        return f
