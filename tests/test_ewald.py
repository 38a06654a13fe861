"""Ewald reciprocal-space forces and energy (frcrec, Sources/bulk.f90:1880,
the VECTOR form the reference build compiles) in the oracle: against golden
vectors written from the reference build (bit-equal there), Newton's third
law, the per-atom energies summing to the total, and the forces being minus
the central-difference gradient of the energy with respect to each atom's
position."""
import os

import numpy as np

import oracle as O

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "reference_vectors.npz"))


def _args(e):
    return (e["ntype"], e["nhk1"], e["nhk2"], e["zv"], e["nel"], e["natom"], e["iatoit"],
            e["iatmpt"], e["ew"], e["tables"])


def test_frcrec_golden(si8):
    import conftest
    e = conftest.ewald_case(si8)
    f, en, ae = O.frcrec(*_args(e), np.full((e["natom"], 3), 0.5), 1.25, np.full(e["natom"], -0.5))
    np.testing.assert_allclose(f, G["frcrec_fclm"], rtol=0, atol=1e-15)
    assert abs(en - float(G["frcrec_ecolmb"])) <= 1e-13 * abs(en)
    np.testing.assert_allclose(ae, G["frcrec_atom_ecolmb"], rtol=1e-13)


def test_frcrec_invariants(si8):
    import conftest
    e = conftest.ewald_case(si8)
    f, en, ae = O.frcrec(*_args(e))
    assert np.abs(f.sum(0)).max() < 1e-13 * np.abs(f).max()
    assert abs(ae.sum() - en) < 1e-12 * abs(en)


def test_frcrec_force_is_energy_gradient(si8):
    import conftest
    e = conftest.ewald_case(si8)
    f, _, _ = O.frcrec(*_args(e))
    L = si8.cell.hcell[0, 0]
    ion = {int(a) - 1: k for k, a in enumerate(e["iatmpt"])}
    h = 1e-4
    for ia in (0, 3):
        for ax in range(3):
            es = []
            for sgn in (1, -1):
                fr = e["frac"].copy()
                fr[ia, ax] += sgn * h / L
                tabs = O.axis_tables(fr[np.asarray(e["iatmpt"]) - 1], e["kcsma"])
                es.append(O.frcrec(*_args(dict(e, tables=tabs)))[1])
            fd = -(es[0] - es[1]) / (2 * h)
            assert abs(f[ia, ax] - fd) < 1e-7 * np.abs(f).max(), (ia, ax, f[ia, ax], fd, ion[ia])
