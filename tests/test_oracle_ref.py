"""Pin the C oracle against the reference's own routines (oracle/_ref, built
with flang from /root/reference by oracle/ref.mk).  Skipped where that library
was not built (the GPU box has no reference tree)."""
import numpy as np
import pytest

import oracle as O
import oracle.ref as R
from qxmd_amd.system import (diamond_si, initial_wavefunctions, model_kb_projectors,
                             occupations)

pytestmark = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("ncell", [1, 2, 3])
def test_gsphere_bitexact(ncell):
    cell = diamond_si(ncell)
    nd = R.get_mesh(cell.hcell, 80.0)
    assert nd == O.get_mesh(cell.hcell, 80.0)
    assert nd == (30 * ncell,) * 3
    s = O.strvec(cell.hcell, 80.0, 20.0)
    r = R.gsphere(cell.hcell, 20.0, 80.0, nd, s["num"] + 1)
    n5 = r["nplw5"]
    for k in ("nplw5", "nplw", "nplw2", "nplwc", "nplw7"):
        assert s[k] == r[k], k
    for k in ("nga", "ngb", "ngc"):
        assert np.array_equal(s[k][: n5 + 1], r[k]), k
    assert np.array_equal(s["recnrm"][: n5 + 1], r["recnrm"])  # bit-exact FP64
    ij = O.ijkgd(s["nga"], s["ngb"], s["ngc"], n5, nd)
    assert np.array_equal(ij, r["ijkgd"])


def test_gsphere_noncubic_cell():
    # an orthorhombic, slightly sheared cell exercises the per-axis kmax and
    # the table rounding of get_mesh for unequal dimensions
    h = np.array([[10.26, 0.3, 0.0], [0.0, 13.1, 0.2], [0.0, 0.0, 8.7]])
    nd = R.get_mesh(h, 80.0)
    assert nd == O.get_mesh(h, 80.0)
    s = O.strvec(h, 80.0, 20.0)
    r = R.gsphere(h, 20.0, 80.0, nd, s["num"] + 1)
    n5 = r["nplw5"]
    assert s["nplw"] == r["nplw"] and s["nplw5"] == n5
    for k in ("nga", "ngb", "ngc"):
        assert np.array_equal(s[k][: n5 + 1], r[k]), k
    assert np.array_equal(O.ijkgd(s["nga"], s["ngb"], s["ngc"], n5, nd), r["ijkgd"])


def test_rnd00():
    a, sa = R.rnd00(25.0, 1000)
    b, sb = O.rnd00_sequence(25.0, 1000)
    assert np.array_equal(a, b) and sa == sb


def _psi(sph, nb, seed=25.0):
    return initial_wavefunctions(sph, nb, seed)


def test_ekinetic(si8):
    psi = _psi(si8, 6)
    occ = occupations(6, 8)
    r1, e1, t1 = O.ekinetic(si8.dkg, psi, occ)
    r2, e2, t2 = R.ekinetic2(psi, occ, si8.dkg)
    assert np.array_equal(r1, r2)
    np.testing.assert_allclose(e1, e2, rtol=1e-14)
    assert abs(t1 - t2) <= 1e-14 * abs(t2)


def test_prpcg(si8):
    for ek in (0.7, 3.1):
        np.testing.assert_allclose(O.prpcg(ek, si8.dkg), R.prpcg(ek, si8.dkg), rtol=1e-15)


def test_hckbpp(si8):
    nb = 5
    psi = _psi(si8, nb)
    rh0 = _psi(si8, nb, seed=77.0)
    eq, lmx, vcl, cl = model_kb_projectors(si8, si8.cell)
    lmx = lmx.copy()
    lmx[1] = 1          # exercise a per-atom projector count
    natom, lmmax = R.kb_setup(eq, lmx, vcl, cl, nb)
    r_ref, sl_ref = R.hckbpp_g(psi, rh0, natom, lmmax)
    r_or, sl_or = O.hckbpp_g(eq, lmx, vcl, psi, rh0)
    np.testing.assert_allclose(r_or, r_ref, rtol=1e-12, atol=1e-14)
    for ia in range(natom):
        np.testing.assert_allclose(sl_or[:, ia, : lmx[ia]], sl_ref[:, ia, : lmx[ia]],
                                   rtol=1e-12, atol=1e-15)
    occ = np.array([2.0, 2.0, 2.0, 0.0, 1.0])
    e_ref = R.enkbpp_g2(nb, natom, occ, si8.cell.volume)
    e_or = O.enkbpp_g(sl_or, lmx, cl, occ, si8.cell.volume)
    assert abs(e_or - e_ref) <= 1e-12 * abs(e_ref)
    # single band entry point hckbpp_ib_g
    r1 = R.hckbpp_ib_g(psi[2], rh0[2], natom)
    np.testing.assert_allclose(r1, r_ref[2], rtol=1e-12, atol=1e-14)


def test_schmidt(si8):
    psi = _psi(si8, 12)
    a, oka = O.schmidt(psi)
    b, okb = R.schmidt2(psi)
    assert oka and okb
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-13)


def test_icholesky():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((40, 25))
    s = x.T @ x + 0.1 * np.eye(25)
    li_ref = R.icholesky(s)
    li = O.rsdecmp_rsinvcd(s)
    np.testing.assert_allclose(np.tril(li), np.tril(li_ref), rtol=1e-12, atol=1e-14)


def test_hmatrix_untrycore(si8):
    psi = _psi(si8, 9)
    h = _psi(si8, 9, seed=5.0)
    a, ea = O.hmatrix(psi, h)
    b, eb = R.hmatrix2(psi, h)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(ea, eb, rtol=1e-13)
    u = np.linalg.qr(np.random.default_rng(1).standard_normal((9, 9)))[0]
    np.testing.assert_allclose(O.untrycore(psi, u), R.untrycore(psi, u), rtol=1e-12, atol=1e-15)


def test_eigh_vs_rseigqr():
    rng = np.random.default_rng(11)
    for n in (7, 40):
        x = rng.standard_normal((n, n))
        a = x + x.T
        w_ref, v_ref = R.rseigqr(a)
        w, v = O.eigh(a)
        np.testing.assert_allclose(w, w_ref, rtol=1e-11, atol=1e-12)
        # eigenvectors agree up to sign (eigenvalues are simple here)
        dots = np.abs(np.sum(v * v_ref, axis=0))
        np.testing.assert_allclose(dots, 1.0, atol=1e-9)


@pytest.mark.parametrize("lclust", [False, True])
def test_stscfn_setnlc_vs_reference(si64, lclust):
    """stscfn_pw (periodic and cluster coordinates) and setnlc_g with one KB
    type, one type skipped (lking) and an imaginary angular row: the oracle
    against the reference routines."""
    import conftest
    c = conftest.nlpp_case(si64, ntype=3)
    ratm = c["frac"] * (si64.cell.hcell.diagonal() if lclust else 1.0)
    args = (ratm, si64.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"], c["kmax"])
    yo, so = O.stscfn_pw(*args, lclust=lclust)
    yr, sr = R.stscfn_pw(*args, lclust=lclust)
    assert np.abs(yo - yr).max() < 1e-14 and np.abs(so - sr).max() < 1e-14
    t = c["tables"]
    t.itolmf[1, 2] = 2  # an angular factor from qlmi
    t.qlmi[2] = 0.5 * t.qlmr[2]
    lking = np.array([0, 1, 0], np.int32)
    eo, vo = O.setnlc_g(t, c["nhk1"], c["nhk2"], c["natom"], yo, so, c["recnrm"],
                        si64.cell.volume, lking=lking)
    er, vr = R.setnlc_g(t, c["nhk1"], c["nhk2"], c["natom"], yo, so, c["recnrm"],
                        si64.cell.volume, lking=lking)
    assert np.abs(eo - er).max() < 1e-14 and np.abs(vo - vr).max() == 0.0
    assert np.abs(er).max() > 0.1


def test_kbnlpp_frc2_vs_reference(si8):
    nb = 7
    psi = initial_wavefunctions(si8, nb, 31.0)
    eq, lmx, vcl, cl = model_kb_projectors(si8, si8.cell)
    occ = np.linspace(2.0, 0.0, nb)
    natom, lmmax = R.kb_setup(eq, lmx, vcl, cl, nb)
    R.hckbpp_g(psi, np.zeros_like(psi), natom, lmmax)
    nw = si8.nplw + 1
    fr = R.kbnlpp_frc2(psi, occ, si8.gx[:nw], si8.gy[:nw], si8.gz[:nw], natom)
    _, sl = O.hckbpp_g(eq, lmx, vcl, psi, np.zeros_like(psi))
    fo = O.kbnlpp_frc2(eq, lmx, vcl, psi, occ, sl, si8.gx[:nw], si8.gy[:nw], si8.gz[:nw])
    assert np.abs(fo - fr).max() < 1e-11 * np.abs(fr).max()


@pytest.mark.parametrize("llk", [[0, 0], [1, 0]])
def test_flocal_g_vs_reference(si64, llk):
    import conftest
    f = conftest.flocal_case(si64, seed=11)
    args = (f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], llk, f["natom"], si64.nga, si64.ngb,
            si64.ngc, si64.gx, si64.gy, si64.gz, f["rhgr"], f["svco"], f["kmax"], f["tables"])
    fo, fr = O.flocal_g(*args), R.flocal_g(*args)
    assert np.array_equal(fo, fr)


def test_xclda_invers_vs_reference():
    rng = np.random.default_rng(2)
    rho = 10 ** rng.uniform(-7, 2, 3000)
    core = np.where(rng.uniform(size=rho.size) < 0.5, 10 ** rng.uniform(-5, 0, rho.size), -1.0)
    for c in (core, -np.ones_like(rho)):
        vo, eo, e4o = O.xclda(rho, c, 0.11)
        vr, er, e4r = R.xclda(rho, c, 0.11)
        assert np.array_equal(vo, vr) and np.array_equal(eo, er) and np.array_equal(e4o, e4r)
    for n in (1, 4, 11):
        x = rng.standard_normal((n, n))
        a = x @ x.T + 0.01 * np.eye(n)
        (wo, io), (wr, ir) = O.invers(a), R.invers(a)
        assert io == ir == 0 and np.array_equal(wo, wr)


def test_frcrec_vs_reference(si64):
    import conftest
    e = conftest.ewald_case(si64, numk2=1500, gamma=0.3)
    args = (e["ntype"], e["nhk1"], e["nhk2"], e["zv"], e["nel"], e["natom"], e["iatoit"],
            e["iatmpt"], e["ew"], e["tables"])
    fo, eo, ao = O.frcrec(*args)
    fr, er, ar = R.frcrec(*args)
    assert np.array_equal(fo, fr) and eo == er and np.array_equal(ao, ar)


def test_roteig_coef_aspc_vs_reference(si8):
    import conftest
    cur, prev = conftest.roteig_case(si8, nb=10, nprv=3, seed=3)
    for kp in (1, 3):
        out, _, rc = O.roteig(cur, prev, kp)
        assert rc == 0
        np.testing.assert_allclose(out, R.roteig(cur, prev, kp), rtol=0, atol=1e-13)
    for ih in (1, 2, 5):
        assert np.array_equal(O.coef_aspc(ih + 1), R.coef_aspc(ih))


@pytest.mark.parametrize("lsreal8", [True, False])
def test_restart_files_vs_reference(si64, tmp_path, lsreal8):
    import conftest
    from qxmd_amd import restart
    from qxmd_amd.system import initial_wavefunctions
    g = initial_wavefunctions(si64, 7, 3.0)
    eig, occ = np.linspace(-1, 1, 7), np.full(7, 2.0)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    restart.svwfns(a, g, eig, occ, lsreal8)
    R.svwfns(b, g, eig, occ, lsreal8)
    assert open(a, "rb").read() == open(b, "rb").read()
    _, _, _, ion = conftest.restart_case(si64)
    ion = dict(ion, lclust=True, ifmd=2)
    restart.svacon(a, **ion)
    R.svacon(b, **ion)
    assert open(a, "rb").read() == open(b, "rb").read()


@pytest.mark.parametrize("seed", [1, 2])
def test_verlet_updates_vs_reference(seed):
    import conftest
    c = conftest.md_case(seed)
    sp, lf = c["species"], c["lfixion"]
    assert np.array_equal(O.updtps(c["x"], c["v"], sp, lf), R.md_kick(0, c["x"], c["v"], sp, lf))
    assert np.array_equal(O.vkick(c["v"], c["f"], sp, lf), R.md_kick(1, c["v"], c["f"], sp, lf))
    assert np.array_equal(O.normalized_a(c["f"], sp, c["hcell"], c["acon"]),
                          R.normalized_a(c["f"], sp, c["hcell"], c["acon"]))
    assert np.array_equal(O.updtps_pbc(c["x"]), R.updtps_pbc(c["x"]))
    b, v = O.rciprl(c["hcell"])
    rb, rv = R.rciprl(c["hcell"])
    assert np.array_equal(b, rb) and v == rv


@pytest.mark.parametrize("which", ["si8", "si64", "sheared"])
def test_clmrcp_bitexact(which):
    """The library's Ewald G list (qx_clmrcp) equals the reference clmrcp's
    (Sources/bulk.f90:556) bit for bit: order, integer vectors, Cartesian G
    and |G|^2, kcsma."""
    from qxmd_amd import pw
    if which == "sheared":
        h = np.array([[10.26, 0.3, 0.0], [0.0, 13.1, 0.2], [0.0, 0.0, 8.7]])
    else:
        h = diamond_si(1 if which == "si8" else 2).hcell
    b = np.linalg.inv(h).T  # reciprocal vectors (columns), without 2 pi
    dkmax2 = 37.0 * 0.35 ** 2 / np.pi ** 2
    r = R.clmrcp(b, dkmax2)
    g = pw.clmrcp(b, dkmax2)
    assert r["numk2"] == len(g["nga"]) > 100
    for k in ("nga", "ngb", "ngc", "gx", "gy", "gz", "recnrm", "kcsma"):
        assert np.array_equal(r[k], g[k]), k
