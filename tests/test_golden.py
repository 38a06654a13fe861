"""Pin the C oracle against golden vectors written from the reference's own
routines (tests/golden/make_golden.py, from the flang build of the reference
sources).  Same inputs and tolerances as tests/test_oracle_ref.py, but these
run wherever the fixtures are, with or without the reference tree."""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from qxmd_amd.system import (diamond_si, initial_wavefunctions, model_kb_projectors,
                             occupations)

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "reference_vectors.npz"))
SHEARED = np.array([[10.26, 0.3, 0.0], [0.0, 13.1, 0.2], [0.0, 0.0, 8.7]])
SUB = 7


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("tag", ["si8", "si64", "sheared"])
def test_golden_gsphere_bitexact(tag):
    h = {"si8": diamond_si(1).hcell, "si64": diamond_si(2).hcell, "sheared": SHEARED}[tag]
    nd = O.get_mesh(h, 80.0)
    assert tuple(G[f"gs_{tag}_nd"]) == tuple(nd)
    s = O.strvec(h, 80.0, 20.0)
    n5 = int(G[f"gs_{tag}_counts"][0])
    assert [s[k] for k in ("nplw5", "nplw", "nplw2", "nplwc", "nplw7")] == list(G[f"gs_{tag}_counts"])
    for k in ("nga", "ngb", "ngc"):
        assert sha(np.asarray(s[k][: n5 + 1], np.int32)) == str(G[f"gs_{tag}_{k}_sha"]), k
    assert sha(np.asarray(s["recnrm"][: n5 + 1], np.float64)) == str(G[f"gs_{tag}_recnrm_sha"])
    ij = O.ijkgd(s["nga"], s["ngb"], s["ngc"], n5, nd)
    assert sha(np.asarray(ij, np.int32)) == str(G[f"gs_{tag}_ijkgd_sha"])


def test_golden_rnd00():
    b, sb = O.rnd00_sequence(25.0, 1000)
    assert np.array_equal(np.asarray(b[:16]), G["rnd00_first"])
    assert sha(np.asarray(b, np.float64)) == str(G["rnd00_sha"])
    assert sb == float(G["rnd00_seed_after"])


def test_golden_ekinetic_prpcg(si8):
    psi = initial_wavefunctions(si8, 6, 25.0)
    _, e1, t1 = O.ekinetic(si8.dkg, psi, occupations(6, 8))
    np.testing.assert_allclose(e1, G["ekinetic_ekib"], rtol=1e-14)
    assert abs(t1 - float(G["ekinetic_ekint"])) <= 1e-14 * abs(float(G["ekinetic_ekint"]))
    for i, ek in enumerate((0.7, 3.1)):
        np.testing.assert_allclose(O.prpcg(ek, si8.dkg)[::SUB], G[f"prpcg_{i}"], rtol=1e-15)


def test_golden_hckbpp(si8):
    nb = 5
    psi = initial_wavefunctions(si8, nb, 25.0)
    rh0 = initial_wavefunctions(si8, nb, 77.0)
    eq, lmx, vcl, cl = model_kb_projectors(si8, si8.cell)
    lmx = lmx.copy()
    lmx[1] = 1
    r_or, sl_or = O.hckbpp_g(eq, lmx, vcl, psi, rh0)
    np.testing.assert_allclose(r_or[:, ::SUB], G["hckbpp_rhcr"], rtol=1e-12, atol=1e-14)
    sl = np.array(sl_or, copy=True)
    for ia in range(sl.shape[1]):
        sl[:, ia, lmx[ia]:] = 0.0
    np.testing.assert_allclose(sl, G["hckbpp_slmir"], rtol=1e-12, atol=1e-15)
    occ = np.array([2.0, 2.0, 2.0, 0.0, 1.0])
    e_or = O.enkbpp_g(sl_or, lmx, cl, occ, si8.cell.volume)
    assert abs(e_or - float(G["enkbpp"])) <= 1e-12 * abs(float(G["enkbpp"]))
    nw = si8.nplw + 1
    f_or = O.kbnlpp_frc2(eq, lmx, vcl, psi, occ, sl_or, si8.gx[:nw], si8.gy[:nw], si8.gz[:nw])
    np.testing.assert_allclose(f_or, G["kbnlpp_frc2"], rtol=1e-11,
                               atol=1e-13 * np.abs(G["kbnlpp_frc2"]).max())
    np.testing.assert_allclose(r_or[2, ::SUB], G["hckbpp_ib"], rtol=1e-12, atol=1e-14)


def test_golden_schmidt_icholesky(si8):
    a, ok = O.schmidt(initial_wavefunctions(si8, 12, 25.0))
    assert ok
    np.testing.assert_allclose(a[:, ::SUB], G["schmidt2"], rtol=1e-10, atol=1e-13)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((40, 25))
    li = O.rsdecmp_rsinvcd(x.T @ x + 0.1 * np.eye(25))
    np.testing.assert_allclose(np.tril(li), G["icholesky"], rtol=1e-12, atol=1e-14)


def test_golden_hmatrix_untrycore(si8):
    p9 = initial_wavefunctions(si8, 9, 25.0)
    h9 = initial_wavefunctions(si8, 9, 5.0)
    a, ea = O.hmatrix(p9, h9)
    np.testing.assert_allclose(a, G["hmatrix2"], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(ea, G["hmatrix2_eig"], rtol=1e-13)
    u = np.linalg.qr(np.random.default_rng(1).standard_normal((9, 9)))[0]
    np.testing.assert_allclose(O.untrycore(p9, u)[:, ::SUB], G["untrycore"], rtol=1e-12,
                               atol=1e-15)


def test_golden_rseigqr():
    rng = np.random.default_rng(11)
    for n in (7, 40):
        xx = rng.standard_normal((n, n))
        w, v = O.eigh(xx + xx.T)
        np.testing.assert_allclose(w, G[f"rseigqr_{n}_w"], rtol=1e-11, atol=1e-12)
        v = v * np.sign(v[np.argmax(np.abs(v), axis=0), np.arange(n)])
        np.testing.assert_allclose(v, G[f"rseigqr_{n}_v"], atol=1e-9)


def test_golden_stscfn_setnlc(si8):
    """Structure factors and the KB table: the oracle against the reference's
    stscfn_pw / setnlc_g (bit-equal here; 1e-14 allowed for a libm with other
    cos/sin roundings), and the structure factors against cos/sin(G . tau)."""
    import conftest
    c = conftest.nlpp_case(si8)
    yc, ys = O.stscfn_pw(c["frac"], si8.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"],
                         c["kmax"])
    np.testing.assert_allclose(yc[:, ::SUB], G["stscfn_ycos"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(ys[:, ::SUB], G["stscfn_ysin"], rtol=0, atol=1e-14)
    gv = 2 * np.pi * np.stack([a[: si8.nplw + 1] for a in c["g"]], 1)
    ph = c["frac"] @ gv.T
    assert np.abs(yc - np.cos(ph)).max() < 1e-13
    assert np.abs(ys - np.sin(ph)).max() < 1e-13
    eq, vcl = O.setnlc_g(c["tables"], c["nhk1"], c["nhk2"], c["natom"], yc, ys, c["recnrm"],
                         si8.cell.volume)
    np.testing.assert_allclose(eq[:, :, ::SUB], G["setnlc_eqlmr"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vcl, G["setnlc_vclalp"], rtol=0, atol=0)


def test_kb_force_is_energy_gradient(si8):
    """kbnlpp_frc2 of the oracle is minus the derivative of the KB energy
    (enkbpp_g2) with respect to the atom position, by central differences of
    the projector table built at displaced positions."""
    from qxmd_amd.system import SiCell
    nb = 5
    psi = initial_wavefunctions(si8, nb, 25.0)
    occ = np.array([2.0, 2.0, 2.0, 0.0, 1.0])
    nw = si8.nplw + 1

    def energy(cell):
        eq, lmx, vcl, cl = model_kb_projectors(si8, cell)
        _, sl = O.hckbpp_g(eq, lmx, vcl, psi, np.zeros_like(psi))
        return O.enkbpp_g(sl, lmx, cl, occ, cell.volume), (eq, lmx, vcl, sl)

    _, (eq, lmx, vcl, sl) = energy(si8.cell)
    f = O.kbnlpp_frc2(eq, lmx, vcl, psi, occ, sl, si8.gx[:nw], si8.gy[:nw], si8.gz[:nw])
    h = 1e-4
    for ia in (0, 5):
        for a in range(3):
            tp, tm = si8.cell.tau.copy(), si8.cell.tau.copy()
            tp[ia, a] += h
            tm[ia, a] -= h
            ep = energy(SiCell(ncell=1, hcell=si8.cell.hcell, tau=tp))[0]
            em = energy(SiCell(ncell=1, hcell=si8.cell.hcell, tau=tm))[0]
            fd = -(ep - em) / (2 * h)
            assert abs(f[ia, a] - fd) < 1e-6 * np.abs(f).max() + 1e-12, (ia, a, f[ia, a], fd)
