"""CPU-side checks of the product library: it loads, exports every symbol the
header declares, and its host set-up (grid rule, G sphere, FFT maps) matches
the oracle bit for bit (which in turn matches the reference)."""
import os

import numpy as np
import pytest

import oracle as O
from qxmd_amd import lib, pw
from qxmd_amd.system import diamond_si


def test_library_exports_header_symbols():
    L = lib.load()
    names = lib.declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n


def test_gemm_schedule_switch():
    """The schedule switch is host state: every mode is taken back, a bad one
    is refused without changing it."""
    assert pw.gemm_schedule() == pw.SCHED_TIMED
    try:
        for m in (pw.SCHED_STREAMK, pw.SCHED_SLICED, pw.SCHED_TIMED):
            pw.set_gemm_schedule(m)
            assert pw.gemm_schedule() == m
        pw.set_gemm_schedule(pw.SCHED_SLICED)
        for bad in (-1, 3):
            with pytest.raises(lib.QxError):
                pw.set_gemm_schedule(bad)
        assert pw.gemm_schedule() == pw.SCHED_SLICED
    finally:
        pw.set_gemm_schedule(pw.SCHED_TIMED)


def test_no_cpu_fallback_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present")
    gs = pw.GSphere(diamond_si(1).hcell)
    with pytest.raises(lib.QxError):
        pw.PlaneWave(gs, 4)


@pytest.mark.parametrize("ncell", [1, 2, 3, 4])
def test_gsphere_bitexact_vs_oracle(ncell):
    cell = diamond_si(ncell)
    g = pw.GSphere(cell.hcell)
    s = O.strvec(cell.hcell, 80.0, 20.0)
    assert g.nd == O.get_mesh(cell.hcell, 80.0) == (30 * ncell,) * 3
    n5 = g.nplw5
    assert (g.nplw, n5) == (s["nplw"], s["nplw5"])
    for k in ("nga", "ngb", "ngc", "recnrm", "gx", "gy", "gz"):
        assert np.array_equal(getattr(g, k)[: n5 + 1], s[k][: n5 + 1]), k
    assert np.array_equal(g.ijkgd(), O.ijkgd(s["nga"], s["ngb"], s["ngc"], n5, g.nd))


def test_plane_wave_counts():
    # exact lattice counts at 20/80 Ry (full sphere = 2*nplw+1)
    expect = {1: (1647, 13133), 2: (13133, 104415), 3: (43819, 351915), 4: (104415, 835537)}
    for n, (wf, dens) in expect.items():
        g = pw.GSphere(diamond_si(n).hcell)
        assert 2 * g.nplw + 1 == wf and 2 * g.nplw5 + 1 == dens


def test_clmrcp_lattice_closed_form():
    """qx_clmrcp's Ewald list is the half of the reciprocal lattice with
    0 < |G|^2 <= dkmax2: every lattice vector in the ball appears once, as
    itself or its negative, never both; |G|^2 = |B n|^2; sorted by |G|^2 up to
    the key's resolution (1e-9 of the largest generated |G|^2)."""
    import itertools
    h = np.array([[10.26, 0.3, 0.0], [0.0, 13.1, 0.2], [0.0, 0.0, 8.7]])
    b = np.linalg.inv(h).T
    dk = 0.5
    g = pw.clmrcp(b, dk)
    n = np.stack([g["nga"], g["ngb"], g["ngc"]], 1)
    got = {tuple(v) for v in n}
    assert len(got) == len(n)
    assert not any(tuple(-v) in got for v in n)
    r = 12
    ball = set()
    for v in itertools.product(range(-r, r + 1), repeat=3):
        if v != (0, 0, 0) and float(np.sum((b @ np.array(v)) ** 2)) <= dk:
            ball.add(v)
    assert all(v in ball for v in got)
    assert {v if v in got else tuple(-np.array(v)) for v in ball} == got
    gv = (b @ n.T).T
    assert np.allclose(g["recnrm"], np.sum(gv ** 2, 1), rtol=1e-14, atol=0)
    assert np.all(np.diff(g["recnrm"]) >= -1e-9 * g["recnrm"].max() * 10)
    assert list(g["kcsma"]) == [int(np.abs(n[:, i]).max()) for i in range(3)]
