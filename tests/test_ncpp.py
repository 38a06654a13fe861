"""The silicon LDA pseudopotential tables (tests/golden/si_ncpp_lda.npz,
written by tests/golden/make_si_ncpp.py from the reference's own set-up
routines) and the host restatements that turn them into per-G arrays
(qxmd_amd/ncpp.py):

* the angular factors equal the reference's qlmr / qlmi for the 8-atom cells
  and are the real spherical harmonics they claim to be (orthonormal on the
  sphere, numpy's Legendre polynomials for m = 0);
* the local form factors equal the reference's svco bit for bit;
* the radial tables of the different cells sample one function (a table of
  one cell interpolated at another's points agrees), and the p projector
  vanishes at q = 0 as a p wave must.
"""
import os

import numpy as np
import pytest

from qxmd_amd import ncpp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "si_ncpp_lda.npz")
CELLS = {"tiny": (1, 6.0, 24.0), "si8": (1, 20.0, 80.0)}


@pytest.fixture(scope="module")
def fx():
    return np.load(GOLD)


def _sphere(tag):
    from conftest import Sphere
    n, e, ed = CELLS[tag]
    return Sphere(n, ecut=e, ecutdens=ed)


@pytest.mark.parametrize("tag", list(CELLS))
def test_angular_factors_match_reference(fx, tag):
    s = _sphere(tag)
    nw = s.nplw + 1
    qr, qi = ncpp.angular_factors(s.gx[:nw], s.gy[:nw], s.gz[:nw], int(fx["lmax"]))
    assert qr.shape == fx[f"{tag}_qlmr"].shape
    assert np.abs(qr - fx[f"{tag}_qlmr"]).max() < 1e-15
    assert np.abs(qi - fx[f"{tag}_qlmi"]).max() < 1e-15


def test_angular_factors_are_spherical_harmonics():
    rng = np.random.default_rng(0)
    v = rng.standard_normal((3, 200000))
    v /= np.linalg.norm(v, axis=0)
    qr, qi = ncpp.angular_factors(v[0], v[1], v[2], 2)
    # with the sqrt(2) of the m > 0 rows, Re and Im rows are real harmonics:
    # orthonormal over the sphere (Monte Carlo average times 4 pi)
    rows = [qr[0], qr[1], qr[2], qi[2], qr[3], qr[4], qi[4], qr[5], qi[5]]
    M = np.array([[4 * np.pi * np.mean(a * b) for b in rows] for a in rows])
    assert np.abs(M - np.eye(9)).max() < 0.02
    # m = 0 rows: s_l sqrt((2l+1)/4pi) P_l(cos t), s = +1, +1, -1 for l = 0, 1, 2
    from numpy.polynomial import legendre
    c = v[2]
    for l, row, sgn in ((0, 0, 1), (1, 1, 1), (2, 3, -1)):
        p = legendre.legval(c, [0] * l + [1])
        np.testing.assert_allclose(qr[row], sgn * np.sqrt((2 * l + 1) / (4 * np.pi)) * p,
                                   atol=1e-14)


@pytest.mark.parametrize("tag", list(CELLS))
def test_form_factors_match_reference_bitwise(fx, tag):
    s = _sphere(tag)
    sp = ncpp.si_lda(fx, tag)
    sv = ncpp.form_factors(s.recnrm[: s.nplw5 + 1], sp["rhocr"], sp["rhocra"], sp["zv"])
    assert np.array_equal(sv, fx[f"{tag}_svco"])


def test_radial_tables_are_one_function(fx):
    """The KB table of each 20 Ry cell, evaluated by setqlm's three-point rule
    at the q points of the Si8 table, agrees with the Si8 table; same for the
    local table."""
    def kb_at(tag, q):
        tb, tba, dl = fx[f"{tag}_tbqlm"], fx[f"{tag}_tbqlma"], float(fx[f"{tag}_dlqlm"])
        m = 2 * np.trunc(q / (2 * dl)).astype(int)
        d = 0.5 * (q / dl - m)
        return d * ((d - 1) * tba[:, m] + tb[:, m + 2] - tb[:, m]) + tb[:, m]

    q = np.arange(0, 540) * float(fx["si8_dlqlm"]) * 0.9
    ref = kb_at("si8", q)
    for tag in ("si64", "si216", "si512"):
        got = kb_at(tag, q)
        assert np.abs(got[:2] - ref[:2]).max() < 1e-5 * np.abs(ref[:2]).max(), tag
    # s and p rows used by the projectors (itola 1 and 2); the p row is 0 at q = 0
    assert sorted(set(fx["itola"][: int(fx["lmx"])])) == [1, 2]
    assert fx["si8_tbqlm"][1, 0] == 0.0 and abs(fx["si8_tbqlm"][0, 0]) > 1.0
    assert int(fx["lmx"]) == 4 and int(fx["lclno"]) == 2 and float(fx["zv"]) == 4.0
