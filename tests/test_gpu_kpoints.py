"""k-point H|psi> on the device (qx_kpoints_set / qx_khpsi_slot), all
(k-point, band) blocks in one launch per stage:

* against the numpy oracle of H_k (oracle/kpoint.py, pinned on the CPU to
  the C oracle's Gamma H) on the 8-atom cell at the eight k-points {0, 1/2}^3
  and two general ones, to 1e-12;
* the per-band launch loop against the batched launch;
* the Gamma-folding identity on the reference's Si LDA pseudopotential: the
  64-atom cell at {0, 1/2}^3 against the Gamma H|psi> (the real-band path,
  pinned to the reference build) of the 512-atom cell built from the
  replicated 64-atom cell, with the replicated local potential.  A Bloch state
  chi of k-point k gives the real supercell states (chi + chi*)/2 and
  (chi - chi*)/2i; the supercell's H on them must equal the same combinations
  of H_k chi.
"""
import os

import numpy as np
import pytest

import oracle as O
from oracle import kpoint as OK
from qxmd_amd import kpoints, ncpp, pw
from qxmd_amd.scf import Scf
from qxmd_amd.system import SiCell, diamond_si

pytestmark = pytest.mark.gpu

NCPP = np.load(os.path.join(os.path.dirname(__file__), "golden", "si_ncpp_lda.npz"))


def _bvec(hcell):
    return 2 * np.pi * np.linalg.inv(hcell).T


def _rand_bands(rng, nb, kin):
    x = rng.standard_normal((nb, len(kin))) + 1j * rng.standard_normal((nb, len(kin)))
    return x / (1.0 + kin)[None, :]


def _setup(cell, tag, kf, nb, seed):
    gs = pw.GSphere(cell.hcell)
    ctx = pw.PlaneWave(gs, 2)
    n1, n2, n3 = gs.nd
    rng = np.random.default_rng(seed)
    V = -0.4 + 0.1 * rng.standard_normal(n1 * n2 * n3)
    ctx.set_glocal(V)
    b = _bvec(cell.hcell)
    npw = ctx.kpoints_set(kf, b, 20.0, nb)
    t, _, _ = ncpp.silicon_cell_inputs(gs, NCPP, tag)
    data = []
    for ik in range(len(kf)):
        hkl, kin = ctx.kpoints_sphere(ik)
        assert len(kin) == npw[ik]
        beta, D = kpoints.kb_projectors(t, cell.tau, np.zeros(cell.natom, int), b, kf[ik], hkl,
                                        cell.volume)
        ctx.kset_kb(ik, beta, D)
        c = _rand_bands(rng, nb, kin)
        ctx.kput(0, ik, c)
        data.append(dict(hkl=hkl, kin=kin, beta=beta[:, 0::2] + 1j * beta[:, 1::2], D=D, c=c))
    return gs, ctx, V.reshape(n3, n2, n1), data


def test_khpsi_vs_oracle_si8():
    kf = np.concatenate([kpoints.gamma_folding_kpoints(2),
                         [[0.25, 0.25, 0.25], [-0.25, 0.125, 0.375]]])
    print("k-point H psi, Si8, 10 k-points x 4 bands", flush=True)
    gs, ctx, V, data = _setup(diamond_si(1), "si8", kf, 4, 1)
    try:
        ctx.khpsi(0, 1)
        ctx.khpsi(0, 2, per_band=True)
        for ik, d in enumerate(data):
            ref = OK.hk_apply(d["c"], d["hkl"], d["kin"], V, d["beta"], d["D"])
            got = ctx.kget(1, ik)
            loop = ctx.kget(2, ik)
            s = np.abs(ref).max()
            assert np.abs(got - ref).max() < 1e-12 * s, ik
            assert np.abs(loop - ref).max() < 1e-12 * s, ik
    finally:
        ctx.close()


def test_khpsi_gamma_folding_si64_si512():
    c64 = diamond_si(2)
    a64 = c64.hcell
    shifts = [np.array([i, j, k]) for i in range(2) for j in range(2) for k in range(2)]
    tau512 = np.concatenate([c64.tau + a64 @ s for s in shifts])
    c512 = SiCell(ncell=4, hcell=2 * a64, tau=tau512)
    kf = kpoints.gamma_folding_kpoints(2)
    nb = 2
    print("k-point H psi, Si64 at {0,1/2}^3 x 2 bands", flush=True)
    gs64, ctx64, V64, data = _setup(c64, "si512", kf, nb, 2)
    try:
        ctx64.khpsi(0, 1)
        y = [ctx64.kget(1, ik) for ik in range(len(kf))]
    finally:
        ctx64.close()
    # the 512-atom cell at Gamma, the reference's tables for it
    print("Gamma H psi, Si512 (replicated Si64), 32 real bands", flush=True)
    gs = pw.GSphere(c512.hcell)
    n1, n2, n3 = gs.nd
    assert (n1, n2, n3) == tuple(2 * v for v in gs64.nd)
    t, svco, zv = ncpp.silicon_cell_inputs(gs, NCPP, "si512")
    nw = gs.nplw + 1
    sup = np.stack([np.asarray(a[:nw]) for a in (gs.nga, gs.ngb, gs.ngc)], 1)
    index = [{tuple(int(v) for v in h): p for p, h in enumerate(d["hkl"])} for d in data]
    kidx, pt = kpoints.supercell_map(sup, 2, index)
    kidm, ptm = kpoints.supercell_map(-sup, 2, index)
    assert np.all(pt >= 0) and np.all(ptm >= 0) and np.array_equal(kidx, kidm)
    phis, want = [], []
    for ik, d in enumerate(data):
        sel = kidx == ik
        for b in range(nb):
            for part in (0, 1):
                chi, hchi = np.zeros(nw, complex), np.zeros(nw, complex)
                chim, hchim = np.zeros(nw, complex), np.zeros(nw, complex)
                chi[sel], chim[sel] = d["c"][b, pt[sel]], d["c"][b, ptm[sel]]
                hchi[sel], hchim[sel] = y[ik][b, pt[sel]], y[ik][b, ptm[sel]]
                if part == 0:
                    phi, hphi = (chi + np.conj(chim)) / 2, (hchi + np.conj(hchim)) / 2
                else:
                    phi, hphi = (chi - np.conj(chim)) / 2j, (hchi - np.conj(hchim)) / 2j
                phis.append(np.stack([phi.real, phi.imag], 1).ravel())
                want.append(np.stack([hphi.real, hphi.imag], 1).ravel())
    phis, want = np.array(phis), np.array(want)
    ctx = pw.PlaneWave(gs, len(phis))
    try:
        frac = np.linalg.solve(c512.hcell, c512.tau.T).T
        sc = Scf(ctx, c512.natom, t, svco, zv)
        sc.set_positions(frac)
        ctx.set_glocal(np.tile(V64, (2, 2, 2)).ravel())
        ctx.put(0, phis)
        ctx.hpsi(0, 1, len(phis))
        got = ctx.get(1, len(phis))
    finally:
        ctx.close()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("Gamma folding: max rel err", err)
    assert err < 1e-11


def test_ksolve_si8_vs_dense():
    """The k-point band solver (block Rayleigh-Ritz, batched rocSOLVER
    generalized eigensolves) at four k-points against numpy's eigvalsh of the
    dense H_k built from its definition (oracle/kpoint.py hk_dense, pinned to
    the band-transform form on the CPU).  The weak random potential leaves the
    free-electron shells nearly degenerate, so 12 bands are solved for and the
    bands below the widest gap among bands 4..10 (which converge at the rate
    that gap sets) are compared: eigenvalues to 1e-8 Ry, residuals
    |H psi - eig psi| small; all 12 Ritz vectors orthonormal."""
    kf = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.25, 0.25, 0.25], [-0.25, 0.5, 0.125]])
    nb = 12
    print("k-point band solver, Si8, 4 k-points x 12 bands", flush=True)
    gs, ctx, V, data = _setup(diamond_si(1), "si512", kf, nb, 3)
    try:
        eig, steps = ctx.ksolve(0, 1, max_iter=100, tol=1e-13)
        print("solver steps", steps, flush=True)
        for ik, d in enumerate(data):
            w = np.linalg.eigvalsh(OK.hk_dense(d["hkl"], d["kin"], V, d["beta"], d["D"]))[:nb]
            g = int(np.argmax(np.diff(w)[3:nb - 2])) + 4
            err = np.abs(eig[ik][:g] - w[:g]).max()
            print("k", ik, "bands", g, "max |eig - dense|", err, flush=True)
            assert err < 1e-8, ik
            psi, hpsi = ctx.kget(0, ik), ctx.kget(1, ik)
            s = np.conj(psi) @ psi.T
            assert np.abs(s - np.eye(nb)).max() < 1e-10
            r = hpsi[:g] - eig[ik][:g, None] * psi[:g]
            assert np.abs(r).max() < 1e-5 * np.abs(hpsi).max()
    finally:
        ctx.close()
