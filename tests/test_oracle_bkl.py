"""Pin the oracle's band-by-band BKL CG (eigen_BKL, Sources/eigen.f90:199-938)
step by step, on a cell small enough for a dense H.

The oracle records every CG iteration (its scalars, the vector x before the
update and the direction h after subgs3/subgs4).  Each quantity is then checked
against what numpy computes from the dense H and the definitions, not against
the oracle's own formulas:

* the new eigenvalue of the BKL branch is the lowest Rayleigh-Ritz value of H
  on span{x, h} (Sources/eigen.f90:593-622);
* x' = alpsq (x + dtnew h) is normalised and <x'|H|x'> equals that value;
* h is P g + beta h_prev projected by classical Gram-Schmidt against the
  already-updated bands (subgs3, coefficients from the unprojected direction)
  and then against x (subgs4; Sources/engrad.f90:1072-1248), with g = eig x -
  H x and P the preconditioner (Sources/eigen.f90:427-452), so a wrong beta
  leaves a residual along h_prev;
* beta = zansap / betk, zansap = <P g|P g>, beta = 0 on the first iteration;
* each exit test (itermx / iteremx with noband, tolera on zansa and on the
  eigenvalue change, roughfzan, roughfeig; Sources/eigen.f90:755-764) fires at
  exactly its threshold: a threshold set just above a traced value stops the
  band at that iteration, just below it does not;
* CG sweeps with Gram-Schmidt and rotation converge to numpy's eigenvalues.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle as O
from qxmd_amd.system import initial_wavefunctions, model_kb_projectors, model_local_potential


def _dense_case(nb=8):
    from conftest import Sphere
    sph = Sphere(1, ecut=6.0, ecutdens=24.0)
    v = model_local_potential(sph, sph.cell)
    eq, lmx, vcl, cl = model_kb_projectors(sph, sph.cell)
    H = O.Hamiltonian(sph.nplw, sph.nd, sph.ijk, v, sph.dkg, eq, lmx, vcl)
    n = sph.nplwex
    basis = np.eye(n)
    basis[1, 1] = 0.0  # Im c(G=0) is not a degree of freedom
    Hm = H.hpsi(basis).T  # Hm[:, i] = H e_i on packed coefficients
    W = np.full(n, 2.0)
    W[0] = 1.0
    psi, _ = O.schmidt(initial_wavefunctions(sph, nb))
    hp = H.hpsi(psi)
    hm, _ = O.hmatrix(psi, hp)
    w, u = O.eigh(hm)
    psi, hp = O.untrycore(psi, u), O.untrycore(hp, u)
    ekib = O.ekinetic(sph.dkg, psi)[1]
    return dict(sph=sph, H=H, Hm=Hm, W=W, psi=psi, hp=hp, eig=w, ekib=ekib, nb=nb)


def _ip(W, a, b):
    return float(a @ (W * b))


@pytest.fixture(scope="module")
def case():
    return _dense_case()


def test_dense_h_is_self_adjoint(case):
    Hm, W = case["Hm"], case["W"]
    keep = np.r_[0, 2:len(W)]
    M = (W[:, None] * Hm)[np.ix_(keep, keep)]
    assert np.abs(M - M.T).max() < 1e-12 * np.abs(M).max()


def test_bkl_steps_against_numpy(case):
    H, Hm, W, nb = case["H"], case["Hm"], case["W"], case["nb"]
    e, g_out, r_out, bz, nr, recs, vecs = H.eigen_bkl_trace(case["eig"], case["psi"], case["hp"],
                                                            case["ekib"], itermx=4, iteremx=4)
    assert len(recs) == int(nr.sum()) > nb
    dkg = case["sph"].dkg
    h_prev = None
    zp_prev = None
    for i, r in enumerate(recs):
        ib = r["ib"] - 1
        x, h = vecs[i]
        # x is normalised and h11 is its Rayleigh quotient
        assert abs(_ip(W, x, x) - 1) < 1e-12
        hx = Hm @ x
        assert abs(r["h11"] - _ip(W, x, hx)) < 1e-11
        # h is orthogonal to x (subgs4 comes last)
        scale = np.sqrt(_ip(W, h, h))
        assert abs(_ip(W, x, h)) < 1e-12 * scale
        # preconditioned gradient, zansap and beta from their definitions
        pg = O.prpcg(case["ekib"][ib], dkg) * (r["h11"] * x - hx)
        zp = _ip(W, pg, pg)
        assert abs(r["zansap"] - zp) < 1e-9 * zp
        if r["iter"] == 1:
            assert r["beta"] == 0.0
            dir0 = pg
        else:
            assert abs(r["beta"] - zp / zp_prev) < 1e-9 * abs(r["beta"])
            dir0 = pg + r["beta"] * h_prev
        # h = subgs4(subgs3(dir0)): classical Gram-Schmidt against the updated
        # lower bands (coefficients from the unprojected direction), then
        # against x; a wrong beta leaves a residual along h_prev
        h1 = dir0 - sum((_ip(W, g_out[j], dir0) * g_out[j] for j in range(ib)), np.zeros_like(x))
        h2 = h1 - _ip(W, x, h1) * x
        assert np.sqrt(_ip(W, h - h2, h - h2)) < 1e-10 * scale
        # the BKL branch: the new eigenvalue is the lowest Ritz value on span{x, h}
        assert r["bkl"] == 1 and r["hhi"] > 0
        V = np.column_stack([x, h])
        A = V.T @ (W[:, None] * (Hm @ V))
        B = V.T @ (W[:, None] * V)
        lam = scipy.linalg.eigh(0.5 * (A + A.T), B, eigvals_only=True)[0]
        assert abs(r["eig"] - lam) < 1e-11 * max(1.0, abs(lam))
        xn = r["alpsq"] * (x + r["dtnew"] * h)
        assert abs(_ip(W, xn, xn) - 1) < 1e-12
        assert abs(_ip(W, xn, Hm @ xn) - r["eig"]) < 1e-11
        gn = r["eig"] * xn - Hm @ xn
        assert abs(r["zansa"] - _ip(W, gn, gn)) < 1e-9 * max(r["zansa"], 1e-20)
        # the next record of this band starts from xn; the last one is the output
        last = i + 1 == len(recs) or recs[i + 1]["ib"] != r["ib"]
        if last:
            assert r["stop"] == 1
            assert np.abs(g_out[ib] - xn).max() < 1e-12
            assert abs(e[ib] - r["eig"]) == 0.0
            assert nr[ib] == r["iter"]
            h_prev = zp_prev = None
        else:
            assert r["stop"] == 0
            assert np.abs(vecs[i + 1][0] - xn).max() < 1e-13
            h_prev, zp_prev = h, zp


def _run(case, **kw):
    return case["H"].eigen_bkl_trace(case["eig"], case["psi"], case["hp"], case["ekib"], **kw)


def _sequences(case, K=10):
    """Per band: zansa after each iteration, the eigenvalue before the first
    and after each iteration (no exit test can fire before K iterations)."""
    *_, recs, _ = _run(case, itermx=K, iteremx=K, tolera=0.0, roughfzan=0.0, roughfeig=0.0,
                       vectors=False)
    seq = {}
    for r in recs:
        s = seq.setdefault(r["ib"] - 1, dict(zansa=[], eig=[r["h11"]]))
        s["zansa"].append(r["zansa"])
        s["eig"].append(r["eig"])
    for s in seq.values():
        assert len(s["zansa"]) == K
    return seq


def _z000re(case, b):
    x = case["psi"][b]
    hx = case["Hm"] @ x
    g = _ip(case["W"], x, hx) * x - hx
    return _ip(case["W"], g, g)


def test_bkl_exit_itermx_noband(case):
    *_, nr, _, _ = _run(case, noband=3, itermx=2, iteremx=5, tolera=0.0, roughfzan=0.0,
                        roughfeig=0.0, vectors=False)
    np.testing.assert_array_equal(nr, [2, 2, 2] + [5] * (case["nb"] - 3))


@pytest.mark.parametrize("which", ["tolera_zansa", "tolera_eig", "roughfzan", "roughfeig"])
def test_bkl_exit_thresholds(case, which):
    K = 10
    seq = _sequences(case, K)
    b = 0  # band 1: no earlier band whose trajectory the threshold could change
    zs, es = seq[b]["zansa"], seq[b]["eig"]
    de = [abs(es[i] - es[i + 1]) for i in range(K)]  # |eigb - eig| after iteration i+1
    z0, x0 = _z000re(case, b), de[0]
    vals = {"tolera_zansa": zs, "tolera_eig": de, "roughfzan": [z / z0 for z in zs],
            "roughfeig": [d / x0 for d in de]}[which]
    kw = (lambda t: dict(tolera=t)) if which.startswith("tolera") else (lambda t: {which: t})

    def first_exit(t):
        """Iteration (1-based) at which band b stops when threshold t is used,
        evaluated on the traced sequences with the rules of eigen.f90:755-764."""
        for it in range(1, K + 1):
            z, d = zs[it - 1], de[it - 1]
            if which.startswith("tolera") and (z < t or d < t):
                return it
            if which == "roughfzan" and z < z0 * t:
                return it
            if which == "roughfeig" and it > 1 and d < x0 * t:
                return it
        return K

    # the first iteration k >= 2 whose traced value alone decides the exit
    ks = [k for k in range(2, K) if first_exit(vals[k - 1] * (1 + 1e-9)) == k
          and first_exit(vals[k - 1] * (1 - 1e-9)) > k]
    assert ks, "no traced value decides an exit"
    k = ks[0]
    above, below = vals[k - 1] * (1 + 1e-9), vals[k - 1] * (1 - 1e-9)
    base = dict(itermx=K, iteremx=K, tolera=0.0, roughfzan=0.0, roughfeig=0.0, vectors=False)
    fa, fb = first_exit(above), first_exit(below)
    for t, expect in ((above, fa), (below, fb)):
        *_, nr, _, _ = _run(case, **{**base, **kw(t)})
        assert nr[b] == expect, (which, t, nr[b], expect)


def test_bkl_sweeps_converge_to_numpy(case):
    """Repeated CG + Gram-Schmidt + rotation reach the lowest eigenvalues of the
    dense H (numpy eigvalsh of the W-symmetrised matrix)."""
    Hm, W, H = case["Hm"], case["W"], case["H"]
    keep = np.r_[0, 2:len(W)]
    sw = np.sqrt(W[keep])
    A = (W[:, None] * Hm)[np.ix_(keep, keep)] / np.outer(sw, sw)
    ref = np.linalg.eigvalsh(0.5 * (A + A.T))
    psi, hp, eig, ekib = case["psi"], case["hp"], case["eig"], case["ekib"]
    for _ in range(40):
        eig, psi, hp, _, _ = H.eigen_bkl(eig, psi, hp, ekib)
        psi, ok = O.schmidt(psi)
        assert ok
        hp = H.hpsi(psi)
        hm, _ = O.hmatrix(psi, hp)
        eig, u = O.eigh(hm)
        psi, hp = O.untrycore(psi, u), O.untrycore(hp, u)
    np.testing.assert_allclose(np.sort(eig)[:7], ref[:7], atol=1e-10)
