"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded inputs.  FP64 stages must agree to round-off (tolerances below
are written per test); the mixed-precision FFT stage to FP32 accuracy."""
import numpy as np
import pytest

import oracle as O
from qxmd_amd import pw
from qxmd_amd.system import (initial_wavefunctions, model_kb_projectors, model_local_potential,
                             occupations)

pytestmark = pytest.mark.gpu


def _setup(sph, nb, kb=True):
    gs = pw.GSphere(sph.cell.hcell)
    assert gs.nplw == sph.nplw and gs.nd == sph.nd
    ctx = pw.PlaneWave(gs, nb)
    v = model_local_potential(sph, sph.cell)
    ctx.set_glocal(v)
    H = None
    if kb:
        eq, lmx, vcl, cl = model_kb_projectors(sph, sph.cell)
        ctx.set_kb(eq, lmx, vcl)
        H = O.Hamiltonian(sph.nplw, sph.nd, sph.ijk, v, sph.dkg, eq, lmx, vcl)
    else:
        H = O.Hamiltonian(sph.nplw, sph.nd, sph.ijk, v, sph.dkg)
    return gs, ctx, H, v


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_slot_roundtrip(si8):
    nb = 3
    gs, ctx, _, _ = _setup(si8, nb, kb=False)
    psi = initial_wavefunctions(si8, nb)
    ctx.put(0, psi)
    back = ctx.get(0, nb)
    assert np.array_equal(back, psi)


@pytest.mark.parametrize("nb", [1, 5])
def test_local_apply_fp64(si8, nb):
    gs, ctx, H, v = _setup(si8, nb, kb=False)
    psi = initial_wavefunctions(si8, nb)
    ctx.put(0, psi)
    ctx.local(0, 1, nb)
    got = ctx.get(1, nb)
    ref = H.hpsi(psi)  # kinetic + local (no KB)
    assert _relerr(got, ref) < 1e-12


def test_local_apply_mixed(si8):
    nb = 6
    gs, ctx, H, v = _setup(si8, nb, kb=False)
    psi = initial_wavefunctions(si8, nb)
    ctx.put(0, psi)
    ctx.local(0, 1, nb, prec=pw.MIXED)
    got = ctx.get(1, nb)
    ref = H.hpsi(psi)
    assert _relerr(got, ref) < 2e-6  # FP32 FFT stage


def test_hpsi_fp64_si64(si64):
    nb = 12
    gs, ctx, H, v = _setup(si64, nb)
    psi = initial_wavefunctions(si64, nb)
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb)
    got = ctx.get(1, nb)
    ref = H.hpsi(psi)
    assert _relerr(got, ref) < 1e-12


def test_hckbpp_dropin(si8):
    nb = 4
    gs, ctx, H, v = _setup(si8, nb)
    eq, lmx, vcl, cl = model_kb_projectors(si8, si8.cell)
    psi = initial_wavefunctions(si8, nb)
    r0 = initial_wavefunctions(si8, nb, seed=9.0)
    got, sl = ctx.hckbpp_g(psi, r0, want_slmir=True, lmmax=eq.shape[1], natom=eq.shape[0])
    ref, slr = O.hckbpp_g(eq, lmx, vcl, psi, r0)
    assert _relerr(got, ref) < 1e-12
    assert _relerr(sl, slr) < 1e-12


def test_hcnonloc_dropin(si8):
    gs, ctx, H, v = _setup(si8, 2, kb=False)
    psi = initial_wavefunctions(si8, 1)[0]
    r0 = initial_wavefunctions(si8, 1, seed=3.0)[0]
    got = ctx.hcnonloc(psi, r0)
    ref = O.hcnonloc(v, psi, r0, si8.nplw, si8.ijk, si8.nd)
    assert _relerr(got, ref) < 1e-12


def test_schmidt(si8):
    nb = 20
    gs, ctx, H, v = _setup(si8, nb, kb=False)
    psi = initial_wavefunctions(si8, nb)
    got, ok = ctx.schmidt_host(psi)
    ref, okr = O.schmidt(psi)
    assert ok and okr
    assert _relerr(got, ref) < 1e-10


def test_schmidt_not_positive_definite(si8):
    nb = 4
    gs, ctx, H, v = _setup(si8, nb, kb=False)
    psi = initial_wavefunctions(si8, nb)
    psi[3] = psi[1]  # linearly dependent set -> lcholesky = .false.
    got, ok = ctx.schmidt_host(psi)
    assert not ok
    assert np.array_equal(got, psi)  # left untouched


def test_rotation(si8):
    nb = 16
    gs, ctx, H, v = _setup(si8, nb)
    psi, _ = O.schmidt(initial_wavefunctions(si8, nb))
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb)
    eig = ctx.rotate(0, 1, nb)
    hp = H.hpsi(psi)
    hm, _ = O.hmatrix(psi, hp)
    w, u = O.eigh(hm)
    np.testing.assert_allclose(eig, w, rtol=0, atol=1e-11)
    p2 = ctx.get(0, nb)
    hp2 = ctx.get(1, nb)
    hm2, e2 = O.hmatrix(p2, hp2)
    assert np.abs(hm2 - np.diag(w)).max() < 1e-10  # rotated basis diagonalises H


def test_density(si8):
    nb = 10
    gs, ctx, H, v = _setup(si8, nb, kb=False)
    psi, _ = O.schmidt(initial_wavefunctions(si8, nb))
    occ = occupations(nb, 16)
    rho = ctx.softchg(psi, occ, si8.cell.volume)
    ref = O.softchg(psi, si8.nplw, occ, si8.cell.volume, si8.ijk, si8.nd)
    assert _relerr(rho, ref) < 1e-12


@pytest.mark.parametrize("nd", [(30, 30, 30), (6, 4, 5), (60, 60, 60)])
def test_rfft3_dropin(nd):
    rng = np.random.default_rng(7)
    N = int(np.prod(nd))
    x = rng.standard_normal(N)
    fx, fy = pw.rfft3(1, x, np.zeros(N), nd)
    rx, ry = O.rfft3(1, x, np.zeros(N), nd)
    assert np.abs(fx - rx).max() < 1e-14 and np.abs(fy - ry).max() < 1e-14
    bx, by = pw.rfft3(2, rx, ry, nd)
    ox, oy = O.rfft3(2, rx, ry, nd)
    assert np.abs(bx - ox).max() < 1e-11 and not by.any()
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    cx, cy = pw.fft3(1, z.real, z.imag, nd)
    qx, qy = O.fft3(1, z.real, z.imag, nd)
    assert np.abs(cx - qx).max() < 1e-14 and np.abs(cy - qy).max() < 1e-14


def test_bkl_band_by_band_fp64(si8):
    nb = 12
    gs, ctx, H, v = _setup(si8, nb)
    psi, _ = O.schmidt(initial_wavefunctions(si8, nb))
    hp = H.hpsi(psi)
    hm, _ = O.hmatrix(psi, hp)
    w, u = O.eigh(hm)
    psi = O.untrycore(psi, u)
    hp = O.untrycore(hp, u)
    ekib = O.ekinetic(si8.dkg, psi)[1]
    e_ref, p_ref, h_ref, bz_ref, nr_ref = H.eigen_bkl(w, psi, hp, ekib)
    ctx.put(0, psi)
    ctx.put(1, hp)
    e, bz, nr = ctx.eigen_bkl(nb, ekib, w)
    np.testing.assert_array_equal(nr, nr_ref)
    np.testing.assert_allclose(e, e_ref, rtol=0, atol=1e-11)
    np.testing.assert_allclose(bz, bz_ref, rtol=1e-7)
    assert _relerr(ctx.get(0, nb), p_ref) < 1e-9
    assert _relerr(ctx.get(1, nb), h_ref) < 1e-9


def test_scf_band_step_fp64(si8):
    nb = 20
    gs, ctx, H, v = _setup(si8, nb)
    psi, _ = O.schmidt(initial_wavefunctions(si8, nb))
    occ = occupations(nb, 32)
    ref = H.scf_band_step(psi, occ, si8.cell.volume)
    ctx.put(0, psi)
    eig, rho, ok = ctx.scf_band_step(nb, occ, si8.cell.volume)
    assert ok and ref["ok"]
    # eigenvalues after CG: Ry -> Ha tolerance 1e-6 Ha per eigenvalue
    np.testing.assert_allclose(eig / 2, ref["eig"] / 2, rtol=0, atol=1e-8)
    assert _relerr(rho, ref["rho"]) < 1e-8


def test_multi_batch_pencil(si64, monkeypatch):
    """Force several band-pair batches through the pencil buffers (1 MiB cap):
    local apply and density must not depend on the batching."""
    monkeypatch.setenv("QX_PENCIL_MB", "1")
    nb = 13
    gs, ctx, H, v = _setup(si64, nb, kb=False)
    psi, _ = O.schmidt(initial_wavefunctions(si64, nb))
    ctx.put(0, psi)
    ctx.local(0, 1, nb)
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < 1e-12
    occ = occupations(nb, 20)
    rho = ctx.density(0, nb, occ, si64.cell.volume)
    ref = O.softchg(psi, si64.nplw, occ, si64.cell.volume, si64.ijk, si64.nd)
    assert _relerr(rho, ref) < 1e-12


@pytest.mark.parametrize("path", ["coop", "legacy"])
@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 129, 200, 540, 1280])
def test_icholesky(si8, n, path, monkeypatch):
    """Cholesky + inverse: the one-launch cooperative kernel and the
    multi-launch fallback against rsdecmp/rsinvcd (n = 1280: the Si512 band
    count)."""
    if path == "legacy":
        monkeypatch.setenv("QX_DENSE_LEGACY", "1")
    gs, ctx, H, v = _setup(si8, max(n, 2), kb=False)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n + 5, n))
    s = x.T @ x + 0.05 * np.eye(n)
    li, ok = ctx.icholesky(s)
    assert ok
    ref = O.rsdecmp_rsinvcd(s)
    assert _relerr(np.tril(li), np.tril(ref)) < 1e-11
    assert np.abs(np.triu(li, 1)).max() == 0.0


def test_icholesky_not_positive_definite(si8):
    gs, ctx, H, v = _setup(si8, 80, kb=False)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((80, 40))
    s = x @ x.T  # rank 40 < 80
    _, ok = ctx.icholesky(s)
    assert not ok


@pytest.mark.parametrize("path", ["coop", "legacy"])
@pytest.mark.parametrize("case", ["random", "degenerate", "clustered"])
def test_rseigqr(si8, case, path, monkeypatch):
    if path == "legacy":
        monkeypatch.setenv("QX_DENSE_LEGACY", "1")
    n = 150
    gs, ctx, H, v = _setup(si8, n, kb=False)
    rng = np.random.default_rng(11)
    q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    if case == "random":
        x = rng.standard_normal((n, n))
        a = x + x.T
    else:
        lam = np.sort(rng.uniform(-3, 3, n))
        if case == "degenerate":
            lam[10:13] = lam[10]       # triply degenerate
            lam[50:52] = lam[50]
        else:
            lam[20:60] = lam[20] + 1e-9 * np.arange(40)  # tight cluster
        a = (q * lam) @ q.T
        a = 0.5 * (a + a.T)
    w, z = ctx.rseigqr(a)
    wr, zr = O.eigh(a)
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-11 * np.abs(wr).max())
    assert np.abs(z.T @ z - np.eye(n)).max() < 1e-10
    assert np.abs(a @ z - z * w).max() < 1e-10 * np.abs(wr).max()


@pytest.mark.parametrize("backtrans", ["wy", "lds"])
@pytest.mark.parametrize("n", [64, 65, 540, 1280])
def test_rseigqr_sizes(si8, n, backtrans, monkeypatch):
    """The eigensolver at the band counts of the silicon cells and around the
    16-row blocks of the back-transformation: the blocked form on MFMA (the
    column panel in registers, one tile-count variant per size range) and the
    reflector-by-reflector form, each against numpy's eigh by residual and
    orthonormality."""
    if backtrans == "lds":
        monkeypatch.setenv("QX_BACKTRANS_LDS", "1")
    gs, ctx, H, v = _setup(si8, n, kb=False)
    rng = np.random.default_rng(n + 7)
    x = rng.standard_normal((n, n))
    a = x + x.T
    w, z = ctx.rseigqr(a)
    wr = np.linalg.eigvalsh(a)
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-11 * np.abs(wr).max())
    assert np.abs(z.T @ z - np.eye(n)).max() < 1e-10
    assert np.abs(a @ z - z * w).max() < 1e-10 * np.abs(wr).max()


@pytest.mark.parametrize("n", [3, 4, 5, 17, 64, 65, 300, 540, 640, 641, 1280])
def test_rseigqr_sizes(si8, n):
    """The cooperative eigensolver over the band counts the workloads use
    (20, 160, 540, 1280), odd sizes around the 64-row tiles, and both sides
    of the switch from four to eight rows per workgroup in the
    tridiagonalisation (n = 640 / 641)."""
    gs, ctx, H, v = _setup(si8, max(n, 4), kb=False)
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal((n, n))
    a = x + x.T + np.diag(np.linspace(0, 5, n))
    w, z = ctx.rseigqr(a)
    wr = np.linalg.eigvalsh(a)
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-11 * np.abs(wr).max())
    assert np.abs(z.T @ z - np.eye(n)).max() < 1e-10
    assert np.abs(a @ z - z * w).max() < 1e-10 * np.abs(wr).max()


def test_generic_fft_path(si64, monkeypatch):
    """The generic Stockham pencil path (used for grids without a two-factor
    split) must agree with the oracle too."""
    monkeypatch.setenv("QX_GENERIC_FFT", "1")
    nb = 7
    gs, ctx, H, v = _setup(si64, nb, kb=False)
    psi, _ = O.schmidt(initial_wavefunctions(si64, nb))
    ctx.put(0, psi)
    ctx.local(0, 1, nb)
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < 1e-12
    occ = occupations(nb, 10)
    rho = ctx.density(0, nb, occ, si64.cell.volume)
    ref = O.softchg(psi, si64.nplw, occ, si64.cell.volume, si64.ijk, si64.nd)
    assert _relerr(rho, ref) < 1e-12


def test_noncubic_cell():
    """Orthorhombic two-cell supercell: grid 30 x 30 x 60 goes through the
    generic pencil path."""
    import conftest
    from qxmd_amd.system import SiCell, diamond_si
    c1 = diamond_si(1)
    tau = np.concatenate([c1.tau, c1.tau + np.array([0.0, 0.0, 10.26])])
    sph = conftest.Sphere(1, cell=SiCell(ncell=1, hcell=np.diag([10.26, 10.26, 20.52]), tau=tau))
    assert sph.nd == (30, 30, 60)
    nb = 5
    gs, ctx, H, v = _setup(sph, nb, kb=True)
    psi = initial_wavefunctions(sph, nb)
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb)
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < 1e-12


@pytest.mark.parametrize("prec", [pw.FP64, pw.MIXED])
def test_fast_path_si216_sample(prec):
    """The 90^3 two-step (9 x 10) path on the 216-atom cell, a few bands."""
    import conftest
    sph = conftest.Sphere(3)
    nb = 6
    gs, ctx, H, v = _setup(sph, nb, kb=False)
    psi = initial_wavefunctions(sph, nb)
    ctx.put(0, psi)
    ctx.local(0, 1, nb, prec=prec)
    tol = 1e-12 if prec == pw.FP64 else 3e-6
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < tol


@pytest.mark.parametrize("prec", [pw.FP64, pw.MIXED])
def test_fast_path_si512_sample(prec):
    """The 120^3 two-step (12 x 10) path on the 512-atom cell, a few bands:
    the FP64 density there keeps six row groups per wavefront."""
    import conftest
    sph = conftest.Sphere(4)
    assert sph.nd == (120, 120, 120)
    nb = 7
    gs, ctx, H, v = _setup(sph, nb, kb=False)
    psi, _ = O.schmidt(initial_wavefunctions(sph, nb))
    ctx.put(0, psi)
    ctx.local(0, 1, nb, prec=prec)
    tol = 1e-12 if prec == pw.FP64 else 3e-6
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < tol
    occ = occupations(nb, 12)
    rho = ctx.density(0, nb, occ, sph.cell.volume, prec=prec)
    ref = O.softchg(psi, sph.nplw, occ, sph.cell.volume, sph.ijk, sph.nd)
    assert _relerr(rho, ref) < tol


def test_hpsi_mixed_si64(si64):
    """Mixed form: FP32 pencil FFT and FP32-MFMA KB contractions against the
    FP64 oracle."""
    nb = 12
    gs, ctx, H, v = _setup(si64, nb)
    psi = initial_wavefunctions(si64, nb)
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb, prec=pw.MIXED)
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < 3e-6  # FP32 FFT + FP32 MFMA


def test_hpsi_mixed_fused_si64(si64, monkeypatch):
    """Mixed band-apply at a size where the KB contractions run between the
    pencil passes (FP32 copy of psi from the z inverse pass, KB term added in
    the z forward pass): the same split bf16x3 products as the plain order,
    whose projections are kept in FP64 rather than FP32, so the two agree to
    FP32 rounding of the KB term; both within the split-product tolerance of
    the FP64 oracle."""
    nb = 256
    gs, ctx, H, v = _setup(si64, nb)
    assert ctx.npj >= 256
    psi = initial_wavefunctions(si64, nb)
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb, prec=pw.MIXED)
    fused = ctx.get(1, nb)
    monkeypatch.setenv("QX_HPSI_UNFUSED", "1")
    ctx.hpsi(0, 1, nb, prec=pw.MIXED)
    plain = ctx.get(1, nb)
    assert _relerr(fused, plain) < 2e-6
    sub = slice(0, 16)
    ref = H.hpsi(psi[sub])
    assert _relerr(fused[sub], ref) < 3e-5
    assert _relerr(plain[sub], ref) < 3e-5


@pytest.mark.parametrize("batched", [False, True])
def test_scf_band_step_mixed_vs_fp64(si64, batched):
    """The mixed SCF band step against the FP64 one from the same start:
    eigenvalues within 1e-6 Ha each, band energy within 1e-6 Ha/atom."""
    nb = 40
    gs, ctx, H, v = _setup(si64, nb)
    psi, _ = O.schmidt(initial_wavefunctions(si64, nb))
    occ = occupations(nb, 64)
    ctx.put(0, psi)
    e64, _, ok64 = ctx.scf_band_step(nb, occ, si64.cell.volume, prec=pw.FP64, batched=batched)
    ctx.put(0, psi)
    emx, _, okmx = ctx.scf_band_step(nb, occ, si64.cell.volume, prec=pw.MIXED, batched=batched)
    assert ok64 and okmx
    # Ry -> Ha
    assert np.abs(np.sort(emx) - np.sort(e64)).max() / 2 < 1e-6
    natom = 64
    assert abs(np.sum(occ * np.sort(emx)) - np.sum(occ * np.sort(e64))) / 2 / natom < 1e-6


@pytest.mark.parametrize("host_sf", [False, True])
def test_stscfn_setnlc_device(si64, host_sf):
    """Structure factors and the KB table built on the device (stscfn_pw,
    setnlc_g) equal the oracle's, which equals the reference's; a band-apply
    with that table matches the oracle Hamiltonian built from the oracle's."""
    import conftest
    c = conftest.nlpp_case(si64, ntype=2)
    t = c["tables"]
    nb = 6
    gs = pw.GSphere(si64.cell.hcell)
    ctx = pw.PlaneWave(gs, nb)
    v = model_local_potential(si64, si64.cell)
    ctx.set_glocal(v)
    ctx.set_nlpp_tables(t)
    yo, so = O.stscfn_pw(c["frac"], si64.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"],
                         c["kmax"])
    yg, sg = ctx.stscfn_pw(c["frac"], si64.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"],
                           c["kmax"])
    np.testing.assert_array_equal(yg, yo)
    np.testing.assert_array_equal(sg, so)
    lking = np.array([0, 1], np.int32) if host_sf else None
    eo, vo = O.setnlc_g(t, c["nhk1"], c["nhk2"], c["natom"], yo, so, c["recnrm"],
                        si64.cell.volume, lking=lking)
    if host_sf:
        vg = ctx.setnlc_g(2, c["nhk1"], c["nhk2"], c["natom"], si64.cell.volume, ycos=yo, ysin=so,
                          lking=lking)
    else:
        vg = ctx.setnlc_g(2, c["nhk1"], c["nhk2"], c["natom"], si64.cell.volume)
    np.testing.assert_array_equal(vg, vo)
    lmmax = t.lmmax
    eg, vclg = ctx.get_kb(c["natom"], lmmax)
    np.testing.assert_array_equal(eg, eo)
    types = np.repeat(np.arange(2), c["natom"] // 2)
    lmx_atom = np.array([0 if (host_sf and types[ia] == 1) else t.lmx[types[ia]]
                         for ia in range(c["natom"])], np.int32)
    vcl_atom = vo[types]
    np.testing.assert_array_equal(vclg, vcl_atom * (lmx_atom[:, None] > 0))
    H = O.Hamiltonian(si64.nplw, si64.nd, si64.ijk, v, si64.dkg, eo, lmx_atom, vcl_atom)
    psi = initial_wavefunctions(si64, nb)
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb)
    assert _relerr(ctx.get(1, nb), H.hpsi(psi)) < 1e-12


def test_kb_forces_device(si64):
    """KB Hellmann-Feynman forces (kbnlpp_frc2) on the device against the oracle."""
    nb = 20
    gs, ctx, H, v = _setup(si64, nb)
    eq, lmx, vcl, cl = model_kb_projectors(si64, si64.cell)
    psi, _ = O.schmidt(initial_wavefunctions(si64, nb))
    occ = np.where(np.arange(nb) < 16, 2.0, 0.0)
    ctx.put(0, psi)
    f0 = np.full((len(lmx), 3), 0.25)
    f = ctx.kb_forces(0, nb, occ, f0)
    nw = si64.nplw + 1
    _, sl = O.hckbpp_g(eq, lmx, vcl, psi, np.zeros_like(psi))
    fo = O.kbnlpp_frc2(eq, lmx, vcl, psi, occ, sl, si64.gx[:nw], si64.gy[:nw], si64.gz[:nw])
    assert np.abs((f - 0.25) - fo).max() < 1e-10 * np.abs(fo).max()


@pytest.mark.parametrize("lking", [False, True])
def test_bkl_band_by_band_device_table(si64, lking, monkeypatch):
    """Band-by-band CG with a KB table built on the device (stscfn_pw +
    setnlc_g): its one-band KB apply runs from the table's factors (phases and
    (type, l) rows) and matches the oracle's CG on the same table, and the
    table-streaming form (QX_KB1_OFF)."""
    import conftest
    c = conftest.nlpp_case(si64, ntype=2)
    t = c["tables"]
    nb = 10
    gs = pw.GSphere(si64.cell.hcell)
    v = model_local_potential(si64, si64.cell)
    lk = np.array([0, 1], np.int32) if lking else None
    yo, so = O.stscfn_pw(c["frac"], si64.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"],
                         c["kmax"])
    eo, vo = O.setnlc_g(t, c["nhk1"], c["nhk2"], c["natom"], yo, so, c["recnrm"],
                        si64.cell.volume, lking=lk)
    types = np.repeat(np.arange(2), c["natom"] // 2)
    lmx_atom = np.array([0 if (lking and types[ia] == 1) else t.lmx[types[ia]]
                         for ia in range(c["natom"])], np.int32)
    H = O.Hamiltonian(si64.nplw, si64.nd, si64.ijk, v, si64.dkg, eo, lmx_atom, vo[types])
    psi, _ = O.schmidt(initial_wavefunctions(si64, nb))
    hp = H.hpsi(psi)
    hm, _ = O.hmatrix(psi, hp)
    w, u = O.eigh(hm)
    psi, hp = O.untrycore(psi, u), O.untrycore(hp, u)
    ekib = O.ekinetic(si64.dkg, psi)[1]
    e_ref, p_ref, h_ref, bz_ref, nr_ref = H.eigen_bkl(w, psi, hp, ekib)
    out = []
    for off in (False, True):
        if off:
            monkeypatch.setenv("QX_KB1_OFF", "1")
        ctx = pw.PlaneWave(gs, nb)
        ctx.set_glocal(v)
        ctx.set_nlpp_tables(t)
        ctx.stscfn_pw(c["frac"], si64.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"],
                      c["kmax"], copy_out=False)
        ctx.setnlc_g(2, c["nhk1"], c["nhk2"], c["natom"], si64.cell.volume, lking=lk)
        ctx.put(0, psi)
        ctx.put(1, hp)
        e, bz, nr = ctx.eigen_bkl(nb, ekib, w)
        np.testing.assert_array_equal(nr, nr_ref)
        np.testing.assert_allclose(e, e_ref, rtol=0, atol=1e-11)
        assert _relerr(ctx.get(0, nb), p_ref) < 1e-9
        assert _relerr(ctx.get(1, nb), h_ref) < 1e-9
        out.append(ctx.get(0, nb))
    assert _relerr(out[0], out[1]) < 1e-12
