"""Device-resident SCF of one cell: the outer loop of pwlda_md
(Sources/pwlda.f90:1628-3700) driven through the C-ABI.

One iteration is the reference's order: the band step (energy pass, subspace
rotation, CG, Gram-Schmidt, density; qx_scf_band_step), outchg (rho(G)),
Pulay mixing (pwpmxchgg) and the new potential V = v_H + v_ext + v_xc
(harting, xclda, pwpunifylc; qx_scf_potential).  The iteration stops when the
Harris-Foulkes energy

    E = sum_n f_n e_n - E_H[rho_in] + E_xc[rho_in] - int v_xc rho_in   (+ core terms)

changes by less than a tolerance per atom; the eigenvalues e_n are those of
the band step.  ``finish`` applies H once more with the last potential and
rotates (hmatrix2 + RsEIGQR + untrycore), so the reported eigenvalues and
energy belong to the final bands and potential.  Per MD step ``set_positions``
rebuilds the structure factors (stscfn_pw), the KB table (setnlc_g) and the
local potential (vlocal_g) on the device.  Forces are the KB
(kbnlpp_frc2), local-pseudopotential (flocal_g) and, when an Ewald G list is
registered, reciprocal-space Ewald (frcrec) terms.

Energies are in Rydberg, as in the reference.
"""
from __future__ import annotations

import numpy as np

from . import pw
from .ncpp import recnrmex


class Scf:
    def __init__(self, ctx: pw.PlaneWave, natom, tables, svco, zv, nhk1=None, nhk2=None,
                 itratn=10, aslh=0.8, bslh=0.6):
        """tables: NlppTables (qx_set_nlpp_tables); svco (ntype, nplw5 + 1)
        local form factors; zv (ntype) valence charges; nhk1 / nhk2 the
        1-based first / last atom of each type (default: one type)."""
        self.ctx, self.gs = ctx, ctx.gs
        self.natom = int(natom)
        self.svco = np.atleast_2d(np.asarray(svco, np.float64))
        self.ntype = self.svco.shape[0]
        self.zv = np.atleast_1d(np.asarray(zv, np.float64))
        self.nhk1 = np.array([1] if nhk1 is None else nhk1, np.int32)
        self.nhk2 = np.array([self.natom] if nhk2 is None else nhk2, np.int32)
        self.aslh, self.bslh = float(aslh), float(bslh)
        gs = self.gs
        n5 = gs.nplw5 + 1
        self._g = [np.asarray(a[:n5], np.int32) for a in (gs.nga, gs.ngb, gs.ngc)]
        self._kmax = [int(np.abs(a).max()) for a in self._g]
        self._nd = np.array(gs.nd, np.int32)
        self.ions = np.arange(1, self.natom + 1, dtype=np.int32)
        self._zero = np.zeros(self.ntype, np.int32)
        ctx.set_nlpp_tables(tables)
        ctx.set_lopp(self.svco)
        ctx.scf_set_grid(recnrmex(gs.recnrm[:n5]))
        ctx.mxchgg_alloc(imxchg=1, itratn=itratn)
        self.ewald = False

    def set_ewald(self, ew):
        self.ctx.set_ewald(ew)
        self.ewald = True

    def set_positions(self, x):
        """Scaled coordinates x (natom, 3): structure factors, KB table and
        local potential at these positions."""
        ctx, gs = self.ctx, self.gs
        ctx.stscfn_pw(np.asarray(x, np.float64), gs.hcell, self._nd, self._nd, *self._g,
                      gs.nplw, self._kmax, copy_out=False)
        ctx.setnlc_g(self.ntype, self.nhk1, self.nhk2, self.natom, gs.volume)
        ctx.vlocal_g(self.ntype, self.nhk1, self.nhk2, self.ions, self._zero, self.natom,
                     gs.volume, copy_out=False)

    def start(self, psi, occ, slot=0):
        """Starting bands into `slot` (and their density as the first input
        density, as the reference starts from the restart file's bands)."""
        ctx = self.ctx
        if psi is not None:
            ctx.put(slot, psi)
        nb = len(occ)
        ctx.pwpchg_slot(slot, nb, occ)
        ctx.scf_put("rinr", ctx.scf_get("rhgr"))
        ctx.pwpmxchgg(self.aslh, self.bslh)
        self.energies = ctx.scf_potential()

    @staticmethod
    def harris_foulkes(occ, eig, en):
        return (float(np.sum(np.asarray(occ) * np.sort(eig))) - en["ehar"] + en["eexc"]
                + en["ecor"] - en["evexc"] - en["evcor"])

    def iterate(self, occ, tol_ha_atom=1e-6, max_scf=60, prec=pw.FP64, batched=True,
                reset_history=True, refine_ha_atom=None):
        """SCF iterations until the Harris-Foulkes energy changes by less than
        tol_ha_atom (Hartree per atom).  Returns (iterations, converged,
        energy history in Ry).

        refine_ha_atom (mixed precision only): once the energy changes by less
        than this, the remaining band steps run in FP64 -- the final FP64
        refinement sweep -- and convergence is only declared on an FP64 step,
        so the converged bands, energy and forces carry FP64 accuracy.
        ``self.fp64_steps`` counts the refinement steps taken."""
        ctx = self.ctx
        nb = len(occ)
        hist = []
        self.fp64_steps = 0
        cur = prec
        for it in range(max_scf):
            eig, _, ok = ctx.scf_band_step(nb, occ, self.gs.volume, prec=cur, batched=batched,
                                           want_rho=False)
            if not ok:
                raise RuntimeError("schmidt: overlap not positive definite")
            if cur == pw.FP64 and prec != pw.FP64:
                self.fp64_steps += 1
            ctx.outchg()
            ctx.pwpmxchgg(self.aslh, self.bslh, lreset=(reset_history and it == 0))
            self.energies = en = ctx.scf_potential()
            hist.append(self.harris_foulkes(occ, eig, en))
            if len(hist) > 1:
                de = abs(hist[-1] - hist[-2]) / 2.0 / self.natom
                if de < tol_ha_atom and (cur == pw.FP64 or refine_ha_atom is None):
                    return it + 1, True, hist
                if refine_ha_atom is not None and cur != pw.FP64 and de < refine_ha_atom:
                    cur = pw.FP64
        return max_scf, False, hist

    def finish(self, occ):
        """Rayleigh-Ritz of the final bands with the final potential (FP64):
        returns (eigenvalues ascending, Harris-Foulkes energy), both in Ry."""
        nb = len(occ)
        self.ctx.hpsi(0, 1, nb, prec=pw.FP64)
        eig = self.ctx.rotate(0, 1, nb)
        return eig, self.harris_foulkes(occ, eig, self.energies)

    def forces(self, occ):
        """Hellmann-Feynman forces (natom, 3), Ry/bohr: KB + local (+ Ewald
        reciprocal space when a G list is registered) on the bands of slot 0
        and the mixed density of the last iteration."""
        ctx = self.ctx
        nb = len(occ)
        f = ctx.kb_forces(0, nb, occ)
        f = f + ctx.flocal_g_dev(self.ntype, self.nhk1, self.nhk2, self.ions, self._zero,
                                 self.natom)
        if self.ewald:
            nel = int(sum((self.nhk2[i] - self.nhk1[i] + 1) * self.zv[i]
                          for i in range(self.ntype)))
            it = np.zeros(self.natom, np.int32)
            for t in range(self.ntype):
                it[self.nhk1[t] - 1: self.nhk2[t]] = t + 1
            fe, _, _ = ctx.frcrec(self.ntype, self.nhk1, self.nhk2, self.zv, nel, self.natom, it,
                                  self.ions)
            f = f + fe
        return f
