"""The batched CG replays one captured iteration as a HIP graph for every
iteration after the first.  Several SCF band steps in a row (each with the
program's own eager work and host synchronisations between the graph
replays: energy pass, rotation, Gram-Schmidt, density) must give the same
bands and eigenvalues, bit for bit, as the same steps with every iteration
launched eagerly."""
import numpy as np
import pytest

from qxmd_amd import pw
from qxmd_amd.system import initial_wavefunctions, model_kb_projectors, model_local_potential

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [pw.FP64, pw.MIXED], ids=["fp64", "mixed"])
def test_cg_graph_replays_match_eager(si64, prec):
    nb = 300  # sizable products: our MFMA kernels run inside the graph
    gs = pw.GSphere(si64.cell.hcell)
    ctx = pw.PlaneWave(gs, nb)
    out = {}
    try:
        ctx.set_glocal(model_local_potential(si64, si64.cell))
        eq, lmx, vcl, _ = model_kb_projectors(si64, si64.cell)
        ctx.set_kb(eq, lmx, vcl)
        ctx.put(0, initial_wavefunctions(si64, nb))
        assert ctx.schmidt(0, nb)
        psi0 = ctx.get(0, nb)
        occ = np.zeros(nb)
        occ[:128] = 2.0
        # one context for both runs: the products' tuned launch shapes are the same
        for graph in (False, True):
            ctx.set_cg_graph(graph)
            ctx.put(0, psi0)
            steps = []
            for k in range(3):
                print("band step", k, "graph" if graph else "eager", flush=True)
                eig, rho, ok = ctx.scf_band_step(nb, occ, si64.cell.volume, prec=prec,
                                                 batched=True)
                assert ok
                steps.append((eig.copy(), ctx.get(0, nb), rho.copy()))
            out[graph] = (steps, ctx.cg_graph_count())
    finally:
        ctx.close()
    eager, graphed = out[False][0], out[True][0]
    assert out[False][1] == 0 and out[True][1] >= 1
    for (e0, p0, r0), (e1, p1, r1) in zip(eager, graphed):
        assert np.array_equal(e0, e1)
        assert np.array_equal(p0, p1)
        assert np.array_equal(r0, r1)
