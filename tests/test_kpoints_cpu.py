"""k-point path on the CPU: the sphere generator (qx_ksphere), the projector
coefficients at k+G (qxmd_amd.kpoints.kb_projectors) and the numpy oracle of
H_k (oracle/kpoint.py), each pinned by something independent of it:

* sphere: brute-force enumeration of |k+G|^2 <= E_cut, and the Gamma
  folding: the spheres of the 8 k-points {0, 1/2}^3 of a cell, mapped to
  G' = 2 (k+G), are exactly the Gamma sphere of its 2 x 2 x 2 supercell from
  the reference's own strvec order (the C oracle, bit-pinned to the
  reference build);
* projectors at k = 0: the Gamma KB table of the oracle (setnlc_g, pinned to
  the reference build) on the half sphere, its conjugate on -G;
* oracle H_k at k = 0 on a real state: the C oracle's Gamma H psi (pinned to
  the reference build); at any k: hermiticity and invariance under relabelling
  k -> k + b_1 (H depends on k + G only).
"""
import itertools

import numpy as np
import pytest

import oracle as O
from oracle import kpoint as OK
from qxmd_amd import kpoints, ncpp, pw
from qxmd_amd.system import diamond_si

NCPP = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                          "si_ncpp_lda.npz"))


def _bvec(hcell):
    return 2 * np.pi * np.linalg.inv(hcell).T


def test_ksphere_brute_force():
    h = np.array([[10.26, 0.4, 0.0], [0.0, 10.9, 0.3], [0.1, 0.0, 9.8]])
    b = _bvec(h)
    k = np.array([0.5, -0.25, 0.125])
    nd = (30, 32, 30)
    hkl, kin = pw.ksphere(b, k, 20.0, nd)
    ref = {}
    for v in itertools.product(range(-20, 21), repeat=3):
        g = (np.array(v) + k) @ b.T
        if g @ g <= 20.0:
            ref[v] = g @ g
    got = [tuple(int(x) for x in r) for r in hkl]
    assert len(got) == len(set(got)) == len(ref)
    assert set(got) == set(ref)
    np.testing.assert_allclose(kin, [ref[v] for v in got], rtol=1e-14, atol=0)
    # storage order: columns (h, k) in lexicographic order, l rising inside
    cols = [(a, bb) for a, bb, _ in got]
    assert cols == sorted(cols)
    for i in range(1, len(got)):
        if cols[i] == cols[i - 1]:
            assert got[i][2] > got[i - 1][2]


def test_ksphere_gamma_folding_union(si64):
    """Spheres of {0, 1/2}^3 on Si8 = the Gamma sphere of the 2 x 2 x 2 cell
    (the replicated Si8 cell has the same lattice as si64)."""
    a8 = np.eye(3) * 10.26
    b8 = _bvec(a8)
    nd8 = O.get_mesh(a8, 80.0)
    union = set()
    total = 0
    for kf in kpoints.gamma_folding_kpoints(2):
        hkl, _ = pw.ksphere(b8, kf, 20.0, nd8)
        total += len(hkl)
        union |= {tuple(int(x) for x in 2 * (r + kf)) for r in hkl}
    half = {(int(a), int(b), int(c)) for a, b, c in
            zip(si64.nga[: si64.nplw + 1], si64.ngb[: si64.nplw + 1], si64.ngc[: si64.nplw + 1])}
    full = half | {(-a, -b, -c) for a, b, c in half}
    assert total == len(union) == len(full) == 2 * si64.nplw + 1
    assert union == full


def _si8_case():
    import conftest
    sp = conftest.Sphere(1)
    t, _, _ = ncpp.silicon_cell_inputs(sp, NCPP, "si8")
    frac = np.linalg.solve(sp.cell.hcell, sp.cell.tau.T).T
    n5 = sp.nplw5 + 1
    g = [np.asarray(a[:n5], np.int32) for a in (sp.nga, sp.ngb, sp.ngc)]
    kmax = [int(np.abs(a).max()) for a in g]
    nd = np.array(sp.nd, np.int32)
    yc, ys = O.stscfn_pw(frac, sp.cell.hcell, nd, nd, *g, sp.nplw, kmax)
    natom = len(frac)
    eq, vcl = O.setnlc_g(t, [1], [natom], natom, yc, ys, np.asarray(sp.recnrm[: sp.nplw + 1]),
                         sp.cell.volume)
    return sp, t, eq, vcl


def test_kb_projectors_gamma_table():
    sp, t, eq, vcl = _si8_case()
    nw = sp.nplw + 1
    hkl = np.stack([np.asarray(a[:nw]) for a in (sp.nga, sp.ngb, sp.ngc)], 1)
    beta, D = kpoints.kb_projectors(t, sp.cell.tau, np.zeros(8, int), _bvec(sp.cell.hcell),
                                    np.zeros(3), hkl, sp.cell.volume)
    lmx = int(t.lmx[0])
    ref = eq[:, :lmx, :].reshape(-1, 2 * nw)  # projector order: atom-major, rows of the type
    scale = np.abs(ref).max()
    assert np.abs(beta - ref).max() < 1e-13 * scale
    np.testing.assert_allclose(D, np.tile(vcl[0, :lmx], 8), rtol=0, atol=0)


def _full_sphere(sp):
    nw = sp.nplw + 1
    h = np.stack([np.asarray(a[:nw]) for a in (sp.nga, sp.ngb, sp.ngc)], 1)
    return np.concatenate([h, -h[1:]]), nw


def test_oracle_hk_gamma_matches_c_oracle():
    """k = 0, a real state (c(-G) = conj c(G)): the numpy H_k equals the C
    oracle's Gamma H psi with the reference's KB table."""
    sp, t, eq, vcl = _si8_case()
    hkl, nw = _full_sphere(sp)
    b = _bvec(sp.cell.hcell)
    kin = np.sum((hkl @ b.T) ** 2, 1)
    rng = np.random.default_rng(5)
    V = -0.4 + 0.1 * rng.standard_normal(sp.N)
    psi = rng.standard_normal((3, 2 * nw)) / (1 + np.asarray(sp.recnrm[:nw]))[None].repeat(2, 1)
    psi[:, 1] = 0.0
    natom = 8
    lmx_atom = np.full(natom, int(t.lmx[0]), np.int32)
    vcl_atom = np.tile(vcl[0], (natom, 1))
    H = O.Hamiltonian(sp.nplw, sp.nd, sp.ijk, V, sp.dkg, eq, lmx_atom, vcl_atom)
    ref = H.hpsi(psi)
    c = psi[:, 0::2] + 1j * psi[:, 1::2]
    cfull = np.concatenate([c, np.conj(c[:, 1:])], 1)
    beta, D = kpoints.kb_projectors(t, sp.cell.tau, np.zeros(natom, int), b, np.zeros(3), hkl,
                                    sp.cell.volume)
    bc = beta[:, 0::2] + 1j * beta[:, 1::2]
    n1, n2, n3 = sp.nd
    got = OK.hk_apply(cfull, hkl, kin, V.reshape(n3, n2, n1), bc, D)[:, :nw]
    refc = ref[:, 0::2] + 1j * ref[:, 1::2]
    assert np.abs(got - refc).max() < 1e-11 * np.abs(refc).max()


def test_oracle_hk_hermitian_and_relabel():
    a = np.eye(3) * 10.26
    b = _bvec(a)
    nd = tuple(int(v) for v in O.get_mesh(a, 80.0))
    n1, n2, n3 = nd
    rng = np.random.default_rng(9)
    V = -0.4 + 0.1 * rng.standard_normal((n3, n2, n1))
    k = np.array([0.25, 0.5, -0.125])
    hkl, kin = pw.ksphere(b, k, 20.0, nd)
    tau = diamond_si(1).tau
    import conftest
    t, _, _ = ncpp.silicon_cell_inputs(conftest.Sphere(1), NCPP, "si8")
    beta, D = kpoints.kb_projectors(t, tau, np.zeros(8, int), b, k, hkl, a[0, 0] ** 3)
    bc = beta[:, 0::2] + 1j * beta[:, 1::2]
    x = rng.standard_normal((2, len(kin))) + 1j * rng.standard_normal((2, len(kin)))
    hx = OK.hk_apply(x, hkl, kin, V, bc, D)
    m = np.conj(x) @ hx.T
    assert np.abs(m - np.conj(m.T)).max() < 1e-12 * np.abs(m).max()
    # k -> k + b1 with labels h -> h - 1: the same vectors k + G, the same H
    k2 = k + np.array([1.0, 0.0, 0.0])
    hkl2 = hkl - np.array([1, 0, 0])
    beta2, _ = kpoints.kb_projectors(t, tau, np.zeros(8, int), b, k2, hkl2, a[0, 0] ** 3)
    hx2 = OK.hk_apply(x, hkl2, kin, V, beta2[:, 0::2] + 1j * beta2[:, 1::2], D)
    assert np.abs(hx2 - hx).max() < 1e-11 * np.abs(hx).max()


def test_oracle_hk_apply_equals_dense_matrix():
    """hk_apply (band transforms) equals hk_dense (the convolution theorem,
    no transforms) on random bands at a general k-point."""
    a = np.eye(3) * 10.26
    b = _bvec(a)
    nd = tuple(int(v) for v in O.get_mesh(a, 80.0))
    n1, n2, n3 = nd
    rng = np.random.default_rng(4)
    V = -0.4 + 0.1 * rng.standard_normal((n3, n2, n1))
    k = np.array([-0.25, 0.5, 0.125])
    hkl, kin = pw.ksphere(b, k, 20.0, nd)
    import conftest
    t, _, _ = ncpp.silicon_cell_inputs(conftest.Sphere(1), NCPP, "si512")
    beta, D = kpoints.kb_projectors(t, diamond_si(1).tau, np.zeros(8, int), b, k, hkl, a[0, 0] ** 3)
    bc = beta[:, 0::2] + 1j * beta[:, 1::2]
    x = rng.standard_normal((3, len(kin))) + 1j * rng.standard_normal((3, len(kin)))
    got = OK.hk_apply(x, hkl, kin, V, bc, D)
    ref = x @ OK.hk_dense(hkl, kin, V, bc, D).T
    assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()
