"""The sizable contractions (M, N >= 256: our 128 x 128 MFMA kernels) through
the C-ABI on the 64-atom cell with 300 bands (npj 256): the KB apply, the
subspace rotation and Gram-Schmidt against the oracle (FP64 to round-off),
the mixed form's split bf16x3 KB products against the FP64 oracle, and our
kernels against hipBLASLt (the library kept as a cross-check only).  K =
nplwex = 13134 is not a multiple of the K stage and 300 bands are not a
multiple of the tile, so the partial stages and edge tiles run."""
import numpy as np
import pytest

import oracle as O
from qxmd_amd import pw
from qxmd_amd.system import initial_wavefunctions, model_kb_projectors, model_local_potential

pytestmark = pytest.mark.gpu

NB = 300


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def case(si64):
    gs = pw.GSphere(si64.cell.hcell)
    ctx = pw.PlaneWave(gs, NB)
    v = model_local_potential(si64, si64.cell)
    ctx.set_glocal(v)
    eq, lmx, vcl, _ = model_kb_projectors(si64, si64.cell)
    ctx.set_kb(eq, lmx, vcl)
    H = O.Hamiltonian(si64.nplw, si64.nd, si64.ijk, v, si64.dkg, eq, lmx, vcl)
    psi, ok = O.schmidt(initial_wavefunctions(si64, NB))
    assert ok
    yield dict(ctx=ctx, H=H, psi=psi, eq=eq, lmx=lmx, vcl=vcl, raw=initial_wavefunctions(si64, NB))
    ctx.close()
    pw.set_gemm_path(pw.GEMM_OWN)


def _kb(ctx, psi, prec=pw.FP64):
    """KB part of H psi alone: hpsi minus the local part."""
    nb = psi.shape[0]
    ctx.put(0, psi)
    ctx.hpsi(0, 1, nb, prec=prec)
    got = ctx.get(1, nb)
    ctx.local(0, 1, nb, prec=prec)
    return got - ctx.get(1, nb)


def test_kb_fp64_vs_oracle(case):
    assert pw.gemm_path() == pw.GEMM_OWN
    ctx = case["ctx"]
    print("KB fp64, 300 bands", flush=True)
    got = _kb(ctx, case["psi"])
    ref, _ = O.hckbpp_g(case["eq"], case["lmx"], case["vcl"], case["psi"],
                        np.zeros_like(case["psi"]))
    err = _relerr(got, ref)
    print("KB fp64 rel err", err)
    assert err < 1e-11


def test_rotation_vs_oracle(case):
    ctx, H, psi = case["ctx"], case["H"], case["psi"]
    print("rotation, 300 bands", flush=True)
    ctx.put(0, psi)
    ctx.hpsi(0, 1, NB)
    eig = ctx.rotate(0, 1, NB)
    hm, _ = O.hmatrix(psi, H.hpsi(psi))
    w, _ = O.eigh(hm)
    np.testing.assert_allclose(eig, w, rtol=0, atol=1e-10)
    hm2, _ = O.hmatrix(ctx.get(0, NB), ctx.get(1, NB))
    assert np.abs(hm2 - np.diag(w)).max() < 1e-9  # the rotated bands diagonalise H


def test_schmidt_vs_oracle(case):
    ctx = case["ctx"]
    print("schmidt, 300 bands", flush=True)
    got, ok = ctx.schmidt_host(case["raw"])
    ref, okr = O.schmidt(case["raw"])
    assert ok and okr
    assert _relerr(got, ref) < 1e-10


@pytest.mark.parametrize("path,tol", [(pw.GEMM_OWN, 1.5e-5), (pw.GEMM_OWN_F32, 2e-6),
                                      (pw.GEMM_LIB, 3e-5)], ids=["bf16x3", "fp32", "hipblaslt"])
def test_kb_mixed_vs_oracle(case, path, tol):
    """Mixed-form KB products (FP32 operands) against the FP64 oracle: split
    bf16x3 products (each factor split into two bf16 parts, about 16
    significant bits, the x_lo y_lo term dropped, FP32 sums; measured 1.05e-5,
    the same as hipBLASLt's fast FP32 mode), plain FP32 MFMA (5e-7), and
    hipBLASLt's fast FP32 mode for comparison."""
    ctx = case["ctx"]
    pw.set_gemm_path(path)
    try:
        print("KB mixed, path", path, flush=True)
        got = _kb(ctx, case["psi"], prec=pw.MIXED)
    finally:
        pw.set_gemm_path(pw.GEMM_OWN)
    ref, _ = O.hckbpp_g(case["eq"], case["lmx"], case["vcl"], case["psi"],
                        np.zeros_like(case["psi"]))
    err = _relerr(got, ref)
    print("KB mixed path", path, "rel err", err)
    assert err < tol


def test_own_kernels_match_hipblaslt(case):
    """FP64 H psi, rotation and Gram-Schmidt with our kernels and with
    hipBLASLt agree to round-off."""
    ctx, psi = case["ctx"], case["psi"]
    out = {}
    for path in (pw.GEMM_OWN, pw.GEMM_LIB):
        pw.set_gemm_path(path)
        try:
            print("fp64 products, path", path, flush=True)
            ctx.put(0, psi)
            ctx.hpsi(0, 1, NB)
            hp = ctx.get(1, NB)
            eig = ctx.rotate(0, 1, NB)
            s, ok = ctx.schmidt_host(case["raw"])
            assert ok
            out[path] = (hp, eig, s)
        finally:
            pw.set_gemm_path(pw.GEMM_OWN)
    a, b = out[pw.GEMM_OWN], out[pw.GEMM_LIB]
    assert _relerr(a[0], b[0]) < 1e-12
    assert np.abs(a[1] - b[1]).max() < 1e-11
    assert _relerr(a[2], b[2]) < 1e-10


@pytest.mark.parametrize("sched", [pw.SCHED_STREAMK, pw.SCHED_SLICED], ids=["streamk", "sliced"])
def test_schedules_vs_oracle(case, sched):
    """Each work schedule of our kernels on its own (stream-K; K-sliced, which
    for the NN products here is one workgroup per output tile): FP64 KB
    apply, split bf16x3 KB apply, rotation and Gram-Schmidt against the
    oracle to the bounds of the tests above."""
    ctx, H, psi = case["ctx"], case["H"], case["psi"]
    pw.set_gemm_schedule(sched)
    try:
        print("schedule", sched, flush=True)
        got64 = _kb(ctx, psi)
        got32 = _kb(ctx, psi, prec=pw.MIXED)
        ctx.put(0, psi)
        ctx.hpsi(0, 1, NB)
        eig = ctx.rotate(0, 1, NB)
        s, ok = ctx.schmidt_host(case["raw"])
        assert ok
    finally:
        pw.set_gemm_schedule(pw.SCHED_TIMED)
    ref, _ = O.hckbpp_g(case["eq"], case["lmx"], case["vcl"], psi, np.zeros_like(psi))
    assert _relerr(got64, ref) < 1e-11
    assert _relerr(got32, ref) < 1.5e-5
    hm, _ = O.hmatrix(psi, H.hpsi(psi))
    w, _ = O.eigh(hm)
    np.testing.assert_allclose(eig, w, rtol=0, atol=1e-10)
    sref, okr = O.schmidt(case["raw"])
    assert okr and _relerr(s, sref) < 1e-10
