"""GPU parity of the wave-function extrapolation between MD steps (roteig2
and sbalinaspc, Sources/sbalin.f90:369-1373) through the C-ABI against the
oracle.  The rotation is unique (orthogonal factor of the overlap), so the
different eigensolvers agree to round-off times the overlap's condition:
1e-11 relative on the turned sets and the predicted bands."""
import numpy as np
import pytest

import oracle as O
from qxmd_amd import pw

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("kpwvx", [1, 2, 3])
def test_roteig_device(si8, kpwvx):
    """Previous sets with large random components (the overlap is far from
    orthogonal, so the polar factor goes through the eigenvalues of A^T A)."""
    import conftest
    cur, prev = conftest.roteig_case(si8, nb=8, nprv=3)
    ctx = pw.PlaneWave(pw.GSphere(si8.cell.hcell), 8)
    ctx.aspc_init(3, 2)
    for k in range(3):
        ctx.aspc_put(k, prev[k], 1)
    ctx.put(0, cur)
    ctx.roteig(0, 8, kpwvx)
    ref, _, rc = O.roteig(cur, prev, kpwvx)
    assert rc == 0
    for k in range(3):
        got = ctx.aspc_get(k, 8)
        assert _rel(got, ref[k]) < 1e-11, (k, _rel(got, ref[k]))
        if k >= kpwvx:
            assert np.array_equal(got, prev[k])


@pytest.mark.parametrize("ncell,nb,ihest", [(1, 8, 2), (1, 8, 3), (2, 24, 3)])
def test_sbalinaspc_device(si8, si64, ncell, nb, ihest):
    """Seven MD steps of a smoothly moving band set: device and oracle state
    machines side by side (iaspc, predicted bands, history)."""
    import conftest
    sp = si8 if ncell == 1 else si64
    cur, prev = conftest.roteig_case(sp, nb=nb, nprv=1)
    d = prev[0] - cur
    n = cur.shape[1]
    st = O.Aspc(nb, n, nprvmx=3, ihest=ihest)
    ctx = pw.PlaneWave(pw.GSphere(sp.cell.hcell), nb)
    ctx.aspc_init(3, ihest)
    for step in range(1, 8):
        g = O.schmidt(cur + 0.05 * step * d + 0.01 * step * step * d[::-1])[0]
        ctx.put(0, g)
        ia_d, ok = ctx.sbalinaspc(0, nb, step)
        pred, ia = st.step(g, step)
        assert ia_d == ia and ok, (step, ia_d, ia)
        assert _rel(ctx.get(0, nb), pred) < 1e-11, (step, _rel(ctx.get(0, nb), pred))
        for k in range(3):
            assert _rel(ctx.aspc_get(k, nb), st.prveig[k]) < 1e-11 or not st.prveig[k].any()


@pytest.mark.parametrize("eigh", [False, True])
def test_roteig_near_orthogonal_overlap(si64, eigh, monkeypatch):
    """Previous sets that are small rotations of the current bands (as between
    MD steps): the polar factor by Newton-Schulz and by the eigenvalues of
    A^T A (QX_ROTEIG_EIGH) agree with the oracle's roteig2."""
    import conftest
    if eigh:
        monkeypatch.setenv("QX_ROTEIG_EIGH", "1")
    nb = 24
    cur, _ = conftest.roteig_case(si64, nb=nb, nprv=1)
    rng = np.random.default_rng(8)
    prev = []
    for k in range(3):
        q, _ = np.linalg.qr(np.eye(nb) + 0.05 * (k + 1) * rng.standard_normal((nb, nb)))
        prev.append(O.schmidt(q.T @ cur + 1e-3 * rng.standard_normal(cur.shape))[0])
    prev = np.stack(prev)
    ctx = pw.PlaneWave(pw.GSphere(si64.cell.hcell), nb)
    ctx.aspc_init(3, 2)
    for k in range(3):
        ctx.aspc_put(k, prev[k], 1)
    ctx.put(0, cur)
    ctx.roteig(0, nb, 3)
    ref, _, rc = O.roteig(cur, prev, 3)
    assert rc == 0
    for k in range(3):
        assert _rel(ctx.aspc_get(k, nb), ref[k]) < 1e-12, k

