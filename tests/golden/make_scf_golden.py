"""Write tests/golden/scf_ncpp_lda.npz: the oracle's converged SCF of the
64-atom silicon cell (2 x 2 x 2 diamond cells,
a = 20.52 bohr, seeded displacements sigma = 0.02 bohr, 20 / 80 Ry, 160
bands, 128 occupied) on the reference's Si LDA pseudopotential tables
(tests/golden/si_ncpp_lda.npz), from the seeded damped random bands.

The CPU oracle needs minutes for this cell, too long for a GPU test, so its
result is kept here: eigenvalues after the final Rayleigh-Ritz step (Ry), the
Harris-Foulkes energy (Ry), the KB + local Hellmann-Feynman forces (Ry/bohr),
the SCF iteration count and the energy history.  The 8-atom cell is included
as well (fast; the GPU tests also run the oracle for it directly).

    python tests/golden/make_scf_golden.py        (needs oracle/liboracle.so)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from oracle.scf import OracleScf  # noqa: E402
from qxmd_amd import ncpp  # noqa: E402
from qxmd_amd.system import initial_wavefunctions, occupations  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "scf_ncpp_lda.npz")
TOL = 1e-9  # Ha/atom between successive SCF energies


def run(tag, ncell, nband):
    from conftest import Sphere
    f = np.load(os.path.join(HERE, "si_ncpp_lda.npz"))
    s = Sphere(ncell)
    sp = ncpp.si_lda(f, tag)
    nw = s.nplw + 1
    t = ncpp.nlpp_tables(s.gx[:nw], s.gy[:nw], s.gz[:nw], [sp])
    svco = ncpp.form_factors(s.recnrm[: s.nplw5 + 1], sp["rhocr"], sp["rhocra"], sp["zv"])
    natom = len(s.cell.tau)
    sc = OracleScf(s, natom, t, svco, [sp["zv"]])
    sc.set_positions(np.linalg.solve(s.cell.hcell, s.cell.tau.T).T)
    occ = occupations(nband, 4 * natom)
    psi, _ = O.schmidt(initial_wavefunctions(s, nband))
    sc.start(psi, occ)
    t0 = time.time()
    it, conv, hist = sc.iterate(occ, tol_ha_atom=TOL, max_scf=200)
    eig, e = sc.finish(occ)
    frc = sc.forces(occ)
    print(tag, "iterations", it, conv, f"{time.time() - t0:.1f} s", "E/atom (Ha)", e / 2 / natom,
          flush=True)
    return {f"{tag}_eig": eig, f"{tag}_energy": np.array(e), f"{tag}_forces": frc,
            f"{tag}_iterations": np.array(it), f"{tag}_converged": np.array(conv),
            f"{tag}_history": np.array(hist)}


def main():
    out = {"tol_ha_atom": np.array(TOL)}
    out.update(run("si8", 1, 20))
    out.update(run("si64", 2, 160))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
