"""SCF glue of the oracle: xclda (Sources/vxc.f90:961), harting
(Sources/engrad.f90:1673), chgr2g/chgg2r (Sources/engrad.f90:1253-1353), INVERS
(Sources/funcs.f90:5) and the Pulay / simple mixers (Sources/chgdns.f90:907-1210).

Pins: xclda and INVERS against golden vectors written from the reference build
(bit-equal there); xclda's potential against the derivative of its energy
density; harting against Poisson's equation for plane-wave densities (the
reference's harting cannot run without FFTW); chgr2g/chgg2r against numpy's
FFT and each other; the Pulay mixer against what Pulay's method fixes (the
coefficients sum to one and minimise the residual in the Kerker metric
computed over the full +-G sphere), its Kerker first step and its history
bookkeeping.  The reference mxchgg_pulay ends in chgg2r -> rfft3 and so cannot
be driven through the reference build either."""
import os

import numpy as np
import pytest

import oracle as O

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "reference_vectors.npz"))


def xc_inputs(n=4000, seed=5):
    """Densities over rs from ~0.1 to ~60 (both Perdew-Zunger branches), a
    zero, a value under the 1e-30 cut-off, and core charge on about 30 % of
    the points (negative rhocore = no core, as the reference tests it)."""
    rng = np.random.default_rng(seed)
    rho = np.concatenate([10 ** rng.uniform(-6, 1.5, n), [0.0, 1e-31, 2.0]])
    core = np.where(rng.uniform(size=rho.size) < 0.3, 10 ** rng.uniform(-4, 0, rho.size), -1.0)
    return rho, core


def test_xclda_golden():
    rho, core = xc_inputs()
    v, e, e4 = O.xclda(rho, core, 0.37)
    np.testing.assert_allclose(v[::7], G["xclda_vexc"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(e[::7], G["xclda_eeexc"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(e4, G["xclda_e4"], rtol=1e-14, atol=0)
    assert v[-3] == 0.0 or core[-3] >= 0  # rho = 0 without core gives zero
    assert v[-2] == 0.0 or core[-2] >= 0


def test_xclda_potential_is_energy_derivative():
    """vexc = d(rho eeexc)/d rho on both Perdew-Zunger branches (rs < 1 and
    rs >= 1), by central differences."""
    rho = np.array([1e-4, 1e-3, 3e-3, 0.01, 0.05, 0.2, 0.3, 1.0, 5.0])
    v, e, _ = O.xclda(rho)
    h = 1e-6 * rho
    _, ep, _ = O.xclda(rho + h)
    _, em, _ = O.xclda(rho - h)
    fd = ((rho + h) * ep - (rho - h) * em) / (2 * h)
    np.testing.assert_allclose(v, fd, rtol=1e-7)
    rs = (3 / (12.56637061436 * rho)) ** (1 / 3)
    assert (rs < 1).any() and (rs > 1).any()
    # exchange alone: eeexc - correlation is -2*0.4582/rs; high density -> exchange dominates
    assert abs(e[-1] / (-2 * 0.4582 / rs[-1]) - 1) < 0.2


def test_xclda_energy_sums(si8):
    rho, core = xc_inputs(500, 9)
    v, e, e4 = O.xclda(rho, core, 0.25)
    rhot = np.where(core >= 0, rho + core, rho)
    on = rhot >= 1e-30
    # eeexc = 2 (ex + ec) integrated over rhot is eexc + ecor
    assert abs(np.sum((e * rhot)[on]) * 0.25 - (e4[0] + e4[2])) < 1e-12 * abs(e4[0] + e4[2])
    # vexc = 2 (vx + vc) integrated over rho is evexc + evcor
    assert abs(np.sum((v * rho)[on]) * 0.25 - (e4[1] + e4[3])) < 1e-12 * abs(e4[1] + e4[3])


def plane_wave_rhgr(sp, ig, re, im):
    rh = np.zeros(2 * (sp.nplw5 + 1))
    rh[2 * ig], rh[2 * ig + 1] = re, im
    return rh


def grid_phase(sp, ig):
    n1, n2, n3 = sp.nd
    i, j, k = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    ph = 2 * np.pi * (sp.nga[ig] * i / n1 + sp.ngb[ig] * j / n2 + sp.ngc[ig] * k / n3)
    return ph.transpose(2, 1, 0).ravel()  # x fastest


def recnrmex(sp):
    r = np.zeros(sp.nplw5 + 1)
    r[1:] = 4 * np.pi / np.asarray(sp.recnrm[1: sp.nplw5 + 1])
    return r


@pytest.mark.parametrize("ig", [1, 7, 500, -1])
def test_harting_poisson_plane_wave(si8, ig):
    """rho(r) = cos(G.r) and sin(G.r): V solves -lap V = 8 pi rho (Rydberg),
    V = 8 pi / G^2 rho, at every grid point."""
    ig = si8.nplw5 if ig == -1 else ig
    rx = recnrmex(si8)
    ph = grid_phase(si8, ig)
    g2 = si8.recnrm[ig]
    for re, im, ref in ((0.5, 0.0, np.cos(ph)), (0.0, -0.5, np.sin(ph))):
        v = O.harting(plane_wave_rhgr(si8, ig, re, im), rx, si8.ijk, si8.nd)
        np.testing.assert_allclose(v, 8 * np.pi / g2 * ref, rtol=0, atol=1e-12 * 8 * np.pi / g2)
        rho = O.chgg2r(plane_wave_rhgr(si8, ig, re, im), si8.ijk, si8.nd)
        np.testing.assert_allclose(rho, ref, rtol=0, atol=1e-12)


def test_harting_g0_and_linearity(si8):
    rng = np.random.default_rng(4)
    rx = recnrmex(si8)
    a = rng.standard_normal(2 * (si8.nplw5 + 1))
    b = rng.standard_normal(2 * (si8.nplw5 + 1))
    a[1] = b[1] = 0.0
    va, vb = O.harting(a, rx, si8.ijk, si8.nd), O.harting(b, rx, si8.ijk, si8.nd)
    vab = O.harting(2 * a - b, rx, si8.ijk, si8.nd)
    np.testing.assert_allclose(vab, 2 * va - vb, atol=1e-10 * np.abs(vab).max())
    # recnrmex(0) = 0 (no double grid): the G = 0 density has no potential
    assert abs(va.mean()) < 1e-12 * np.abs(va).max()


def test_chgr2g_chgg2r_numpy(si8):
    """chgr2g = numpy's forward FFT / N gathered on the sphere, chgg2r its
    inverse for sphere-limited densities."""
    rng = np.random.default_rng(2)
    n1, n2, n3 = si8.nd
    rho = rng.uniform(0.0, 0.1, si8.N)
    rh = O.chgr2g(rho, si8.nplw5, si8.ijk, si8.nd)
    f = np.fft.fftn(rho.reshape(n3, n2, n1)).ravel() / si8.N
    idx = np.asarray(si8.ijk[si8.nplw5: 2 * si8.nplw5 + 1]) - 1
    np.testing.assert_allclose(rh[0::2], f[idx].real, atol=1e-15)
    np.testing.assert_allclose(rh[1::2], f[idx].imag, atol=1e-15)
    back = O.chgg2r(rh, si8.ijk, si8.nd)
    again = O.chgr2g(back, si8.nplw5, si8.ijk, si8.nd)
    np.testing.assert_allclose(again, rh, atol=1e-15)
    assert abs(rh[0] - rho.mean()) < 1e-15


def test_invers_golden():
    rng = np.random.default_rng(21)
    for n in (2, 5, 9):
        x = rng.standard_normal((n, n))
        wi, ierr = O.invers(x @ x.T + 0.05 * np.eye(n))
        assert ierr == 0
        np.testing.assert_allclose(wi, G[f"invers_{n}"], rtol=0, atol=0)
    _, ierr = O.invers(np.ones((3, 3)))
    assert ierr == 1


def full_metric(sp, nplwcex, factq1, rvol):
    """Weights of the Pulay metric over the real-packed vector, derived from
    the full +-G sphere: each G != 0 enters as +G and -G (factor 2 on Re^2 +
    Im^2) with (G^2 + q1)/G^2; G = 0 is real and takes the weight of the first
    shell (the reference evaluates the weight at packed index 3 for it)."""
    nw = nplwcex // 2
    g2 = np.asarray(sp.recnrm[:nw])
    w = np.zeros(nplwcex)
    w[2:] = 2.0 * np.repeat((g2[1:] + factq1) / g2[1:], 2)
    w[0] = (g2[1] + factq1) / g2[1]
    return w * rvol


@pytest.mark.parametrize("lmixhold", [False, True])
def test_pulay_mixer_properties(si8, lmixhold):
    """itratn = 5: without lmixhold the history restarts with a Kerker step
    after 5 entries (kntchg 1..5, 1..); with lmixhold it slides (kntchg stays
    at 4 after the shift and the oldest entry drops out)."""
    rng = np.random.default_rng(8)
    nw = si8.nplwc + 1
    ncx, n5x = 2 * nw, 2 * (si8.nplw5 + 1)
    dkg = O.dkgnrm(si8.recnrm, si8.nplwc)[:ncx]
    q1 = O.factq1(20.0, si8.recnrm[1], si8.recnrm[si8.nplwc])
    assert q1 > 0
    rvol = si8.cell.volume
    itr = 5
    m = O.PulayMixer(ncx, itr, dkg, rvol, q1)
    w = full_metric(si8, ncx, q1, rvol)
    rin = rng.standard_normal(n5x) * 0.01
    rin[1] = 0.0
    hist = []
    for it in range(12):
        rout = rin + rng.standard_normal(n5x) * 0.01 / (1 + it)
        rout[1] = 0.0
        mixed, rinr = m.mix(rout, rin, 0.8, 0.6, lmixhold=lmixhold)
        fnew = (rout - rin)[:ncx].copy()
        if m.kntchg == 1:
            f = 0.8 * (dkg / (dkg + 0.6))
            np.testing.assert_allclose(mixed[:ncx], rin[:ncx] + f * fnew, rtol=1e-15)
            assert not m.cmk.any()
            hist = [fnew]
        else:
            hist.append(fnew)
            n = len(hist)
            c = m.cmk[:n]
            assert abs(c.sum() - 1.0) < 1e-12
            # Pulay: sum c_i F_i has the same metric overlap with every F_i
            F = np.array(hist)
            r = c @ F
            ov = F @ (w * r)
            assert np.ptp(ov) < 1e-9 * np.abs(ov).max(), ov
            # and the combined residual is no larger than the newest one's
            assert r @ (w * r) <= F[-1] @ (w * F[-1]) * (1 + 1e-12)
            if n == itr:
                hist = hist[1:] if lmixhold else hist
        if not lmixhold:
            assert m.kntchg == (it % itr) + 1, (it, m.kntchg)
        else:
            assert m.kntchg == (it + 1 if it + 1 < itr else itr - 1), (it, m.kntchg)
        np.testing.assert_array_equal(mixed[ncx:], rout[ncx:])
        np.testing.assert_array_equal(rinr, mixed)
        rin = mixed
    mixed, _ = m.mix(rin + 0.001, rin, 0.5, 0.6, lreset=True)
    assert m.kntchg == 1


def test_simple_mixers():
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(50), rng.standard_normal(50)
    rh, ri = O.mxchgg_simple(a, b, 0.3)
    np.testing.assert_allclose(rh, 0.7 * b + 0.3 * a, rtol=1e-15)
    np.testing.assert_array_equal(ri, rh)
    rh, ri = O.mxchgg_simple(a, b, 0.3, none=True)
    np.testing.assert_array_equal(rh, a)
    np.testing.assert_array_equal(ri, a)


@pytest.mark.parametrize("tag,llk", [("", [0, 0]), ("_lking", [0, 1])])
def test_flocal_g_golden(si8, tag, llk):
    """flocal_g against the reference build's output (bit-equal there): a
    permuted iatmpt, a skipped (llking) type, and the incoming floc doubled
    with the sum as the reference does."""
    import conftest
    f = conftest.flocal_case(si8)
    out = O.flocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], llk, f["natom"], si8.nga,
                     si8.ngb, si8.ngc, si8.gx, si8.gy, si8.gz, f["rhgr"], f["svco"], f["kmax"],
                     f["tables"], np.full((f["natom"], 3), 0.25))
    np.testing.assert_allclose(out, G["flocal_g" + tag], rtol=0, atol=1e-13)
    if llk[1]:
        skipped = np.asarray(f["iatmpt"])[f["nhk1"][1] - 1: f["nhk2"][1]] - 1
        assert np.all(out[skipped] == 0.25)


def test_flocal_g_is_energy_gradient(si8):
    """floc = -dE/dtau for E = 2 sum_{G != 0} svco(G) Re(rho(G) exp(i G.tau))
    (the local energy over the +-G sphere), phases from numpy cos/sin of
    2 pi n.s, central differences in the Cartesian position (cubic cell)."""
    import conftest
    f = conftest.flocal_case(si8)
    n5 = si8.nplw5
    L = si8.cell.hcell[0, 0]
    ng = np.stack([np.asarray(a[: n5 + 1]) for a in (si8.nga, si8.ngb, si8.ngc)], 1)
    a, b = f["rhgr"][0::2], f["rhgr"][1::2]
    types = np.zeros(f["natom"], int)
    for it in range(2):
        types[np.asarray(f["iatmpt"])[f["nhk1"][it] - 1: f["nhk2"][it]] - 1] = it

    def energy(ia, s):
        ph = 2 * np.pi * (ng[1:] @ s)
        sv = f["svco"][types[ia], 1:]
        return 2 * np.sum(sv * (a[1:] * np.cos(ph) - b[1:] * np.sin(ph)))

    fl = O.flocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], [0, 0], f["natom"], si8.nga,
                    si8.ngb, si8.ngc, si8.gx, si8.gy, si8.gz, f["rhgr"], f["svco"], f["kmax"],
                    f["tables"])
    h = 1e-5
    for ia in (0, 5):
        for ax in range(3):
            sp_, sm = f["frac"][ia].copy(), f["frac"][ia].copy()
            sp_[ax] += h / L
            sm[ax] -= h / L
            fd = -(energy(ia, sp_) - energy(ia, sm)) / (2 * h)
            assert abs(fl[ia, ax] - fd) < 1e-6 * np.abs(fl).max(), (ia, ax, fl[ia, ax], fd)
