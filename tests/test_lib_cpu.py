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
