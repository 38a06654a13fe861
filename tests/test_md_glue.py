"""Wave-function extrapolation between MD steps in the oracle: roteig2
(Sources/sbalin.f90:772), set_coef_ASPC (77) and sbalinaspc (369).

Pins: roteig and the ASPC coefficients against golden vectors of the
reference build (roteig to 1e-13: the oracle's eigensolver is not the
reference's, the rotation it builds is unique); roteig's rotation against the
orthogonal Procrustes solution from numpy's SVD; the coefficients against
Kolafa's published values and their consistency conditions; the composed
sbalinaspc against what extrapolation must do on a fixed subspace (return the
current bands) and on a moving one (predict better than the last step)."""
import os

import numpy as np
import pytest

import oracle as O

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "reference_vectors.npz"))
SUB = 7


def packed(x, y):
    """<x|y> of real-packed half-sphere vectors: 2 sum - x0 y0."""
    return 2 * x @ y.T - np.outer(x[:, 0], y[:, 0])


def test_roteig_golden_and_procrustes(si8):
    import conftest
    cur, prev = conftest.roteig_case(si8)
    out, U, rc = O.roteig(cur, prev, 2)
    assert rc == 0
    np.testing.assert_allclose(out[:, :, ::SUB], G["roteig_prev"], rtol=0, atol=1e-13)
    np.testing.assert_array_equal(out[2], prev[2])  # only kpwvx = 2 sets turned
    # U is the orthogonal factor of A^T, A(i, j) = <prev_i|cur_j>
    A = packed(prev[0], cur)
    P, _, Qt = np.linalg.svd(A.T)
    np.testing.assert_allclose(U, P @ Qt, atol=1e-12)
    np.testing.assert_allclose(U @ U.T, np.eye(len(U)), atol=1e-12)
    # the turned set is the closest rotation of prev to cur
    d_new = np.linalg.norm(out[0] - cur)
    rng = np.random.default_rng(0)
    for _ in range(5):
        Q = np.linalg.qr(U + 0.01 * rng.standard_normal(U.shape))[0]
        assert np.linalg.norm(Q @ prev[0] - cur) >= d_new * (1 - 1e-12)


@pytest.mark.parametrize("ihest", [1, 2, 3, 4])
def test_coef_aspc(ihest):
    c = O.coef_aspc(ihest + 1)
    np.testing.assert_array_equal(c, G[f"coef_aspc_{ihest}"])
    for k in range(2, ihest + 2):
        col = c[:k, k - 1]
        assert abs(col.sum() - 1) < 1e-14  # constant trajectories are kept
        assert abs(col @ (1 - np.arange(1, k + 1)) - 1) < 1e-13  # linear ones extrapolated
    np.testing.assert_allclose(O.coef_aspc(3)[:, 2], [2.5, -2.0, 0.5], rtol=1e-15)
    np.testing.assert_allclose(O.coef_aspc(2)[:, 1], [2.0, -1.0], rtol=1e-15)


def test_sbalinaspc_fixed_subspace(si8):
    """Bands that only rotate within one subspace: once aligned, every
    previous set equals the current one and the prediction is the current
    bands (the coefficients sum to one)."""
    import conftest
    cur, _ = conftest.roteig_case(si8, nprv=1)
    nb, n = cur.shape
    st = O.Aspc(nb, n, nprvmx=3, ihest=3)
    rng = np.random.default_rng(4)
    for step in range(1, 7):
        Q = np.linalg.qr(rng.standard_normal((nb, nb)))[0]
        g = Q @ cur
        pred, ia = st.step(g, step)
        if step == 1:
            assert ia == 0 and np.array_equal(pred, g)
        else:
            assert ia == min(step, 4)
            np.testing.assert_allclose(pred, g, atol=1e-12)


def test_sbalinaspc_moving_subspace(si8):
    """Bands moving along a smooth path: the ASPC prediction is closer to the
    next step's subspace than the current bands are."""
    import conftest
    cur, prev = conftest.roteig_case(si8, nprv=1)
    nb, n = cur.shape
    d = prev[0] - cur
    st = O.Aspc(nb, n, nprvmx=2, ihest=2)

    def path(t):
        return O.schmidt(cur + 0.05 * t * d + 0.01 * t * t * d[::-1])[0]

    def dist(a, b):  # distance of the spans
        s = np.linalg.svd(packed(a, b), compute_uv=False)
        return np.sqrt(np.sum(1 - np.minimum(s, 1) ** 2))

    for step in range(1, 6):
        g = path(step)
        pred, ia = st.step(g, step)
        if step >= 3:
            assert dist(pred, path(step + 1)) < 0.2 * dist(g, path(step + 1)), step
        np.testing.assert_allclose(packed(pred, pred), np.eye(nb), atol=1e-12)
