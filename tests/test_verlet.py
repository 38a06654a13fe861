"""Velocity-Verlet updates of the MD loop (normalized_a, vkick, updtps,
updtps_pbc, rciprl; Sources/remd.f90, Sources/reclat.f90) through the C-ABI
(host arithmetic: no GPU needed) and in the oracle.

Pins: the oracle and the library against golden vectors written by the
reference build, bit for bit; rciprl against numpy's inverse (a_i . b_j =
delta_ij) and determinant; the composed integrator against what velocity
Verlet must do: exact time reversal of a trajectory when the velocities are
negated, and energy conservation of a harmonic crystal to O(dt^2)."""
import os

import numpy as np
import pytest

import conftest
import oracle as O
from qxmd_amd import md
from qxmd_amd.lib import QxError

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "reference_vectors.npz"))


def test_oracle_against_reference_vectors():
    c = conftest.md_case()
    assert np.array_equal(O.updtps(c["x"], c["v"], c["species"], c["lfixion"]), G["md_updtps"])
    assert np.array_equal(O.vkick(c["v"], c["f"], c["species"], c["lfixion"]), G["md_vkick"])
    assert np.array_equal(O.normalized_a(c["f"], c["species"], c["hcell"], c["acon"]),
                          G["md_normalized_a"])
    assert np.array_equal(O.updtps_pbc(c["x"]), G["md_updtps_pbc"])
    b, v = O.rciprl(c["hcell"])
    assert np.array_equal(b, G["md_rciprl_b"]) and v == float(G["md_rciprl_v"])


def test_library_against_reference_vectors():
    c = conftest.md_case()
    assert np.array_equal(md.updtps(c["x"], c["v"], c["species"], c["lfixion"]), G["md_updtps"])
    assert np.array_equal(md.vkick(c["v"], c["f"], c["species"], c["lfixion"]), G["md_vkick"])
    assert np.array_equal(md.normalized_a(c["f"], c["species"], c["hcell"], c["acon"]),
                          G["md_normalized_a"])
    x = md.updtps_pbc(c["x"])
    assert np.array_equal(x, G["md_updtps_pbc"])
    assert np.all((x >= 0.0) & (x < 1.0))
    b, v = md.rciprl(c["hcell"])
    assert np.array_equal(b, G["md_rciprl_b"]) and v == float(G["md_rciprl_v"])


def test_fixed_species_and_edges():
    c = conftest.md_case()
    out = md.updtps(c["x"], c["v"], c["species"], c["lfixion"])
    fixed = c["species"] == 2
    assert np.array_equal(out[fixed], c["x"][fixed])
    assert np.array_equal(out[~fixed], c["x"][~fixed] + c["v"][~fixed])
    edges = md.updtps_pbc(np.array([[0.0, -0.0, 1.0], [-1.0, 2.0, -2.5]]))
    assert edges.tolist() == [[0.0, -0.0, 0.0], [0.0, 0.0, 0.5]]
    with pytest.raises(QxError):
        md.updtps_pbc(np.array([[3e9, 0.0, 0.0]]))
    with pytest.raises(QxError):
        md.vkick(c["v"], c["f"], np.where(c["species"] == 3, 4, c["species"]), c["lfixion"])


def test_rciprl_closed_form():
    c = conftest.md_case()
    h = c["hcell"]
    b, v = md.rciprl(h)
    np.testing.assert_allclose(h.T @ b, np.eye(3), atol=1e-15)
    np.testing.assert_allclose(b, np.linalg.inv(h).T, rtol=1e-14)
    assert v == pytest.approx(abs(np.linalg.det(h)), rel=1e-14)


def _harmonic(x, x0, hcell, k):
    """Cartesian forces -k (r - r0) of a crystal of independent springs."""
    return -k * ((x - x0) @ hcell.T)


def _run(x, v, nsteps, x0, hcell, k, species, acon, lfix):
    a = md.normalized_a(_harmonic(x, x0, hcell, k), species, hcell, acon)
    for _ in range(nsteps):
        v = md.vkick(v, a, species, lfix)
        x = md.updtps_pbc(md.updtps(x, v, species, lfix))
        a = md.normalized_a(_harmonic(x, x0, hcell, k), species, hcell, acon)
        v = md.vkick(v, a, species, lfix)
    return x, v


def test_verlet_reversible_and_conserving():
    c = conftest.md_case()
    rng = np.random.default_rng(3)
    n = len(c["species"])
    hcell, species = c["hcell"], c["species"]
    lfix = np.zeros(3, bool)
    mass = np.array([51196.0, 21874.0, 1837.0])
    dt, k = 40.0, 0.2
    acon = md.acon(mass, dt)
    np.testing.assert_array_equal(acon, c["acon"])
    x0 = rng.uniform(0.3, 0.7, (n, 3))
    x = x0 + rng.normal(0.0, 0.01, (n, 3))
    v = np.zeros((n, 3))
    m = mass[species - 1][:, None]

    def energy(x, v):
        r = (x - x0) @ hcell.T
        vc = (v @ hcell.T) / dt
        return 0.5 * np.sum(m * vc * vc) + 0.5 * k * np.sum(r * r)

    e0 = energy(x, v)
    x1, v1 = _run(x, v, 200, x0, hcell, k, species, acon, lfix)
    e1 = energy(x1, v1)
    # period 2 pi sqrt(m / k) >= 600 au for the lightest species: dt = 40
    # resolves it, and the Verlet energy error stays O((omega dt)^2)
    assert abs(e1 - e0) < 2e-2 * abs(e0)
    assert np.abs(x1 - x).max() > 1e-3  # the atoms did move
    xb, vb = _run(x1, -v1, 200, x0, hcell, k, species, acon, lfix)
    np.testing.assert_allclose(xb, x, rtol=0, atol=1e-13)
    np.testing.assert_allclose(-vb, v, rtol=0, atol=1e-13)
