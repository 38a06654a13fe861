"""Write tests/golden/reference_vectors.npz: outputs of the reference's own
routines (the flang build of the reference sources in oracle/_ref, made by
oracle/ref.mk) on the deterministic inputs that tests/test_golden.py rebuilds.
With these vectors the oracle stays pinned where the reference tree is absent.

    python tests/golden/make_golden.py      (needs oracle/_ref built)
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
import oracle.ref as R  # noqa: E402
from qxmd_amd.system import (diamond_si, initial_wavefunctions, model_kb_projectors,  # noqa: E402
                             occupations)

SHEARED = np.array([[10.26, 0.3, 0.0], [0.0, 13.1, 0.2], [0.0, 0.0, 8.7]])
SUB = 7  # stride of the subsample kept of long vectors


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not R.available():
        raise SystemExit("oracle/_ref is not built (make -f oracle/ref.mk)")
    import conftest
    out = {}
    # G sphere, grid, index maps: bit-exact digests
    for tag, h in (("si8", diamond_si(1).hcell), ("si64", diamond_si(2).hcell), ("sheared", SHEARED)):
        nd = R.get_mesh(h, 80.0)
        s = O.strvec(h, 80.0, 20.0)
        r = R.gsphere(h, 20.0, 80.0, nd, s["num"] + 1)
        out[f"gs_{tag}_nd"] = np.array(nd, np.int32)
        out[f"gs_{tag}_counts"] = np.array([r[k] for k in ("nplw5", "nplw", "nplw2", "nplwc", "nplw7")],
                                           np.int64)
        for k in ("nga", "ngb", "ngc"):
            out[f"gs_{tag}_{k}_sha"] = np.array(sha(np.asarray(r[k], np.int32)))
        out[f"gs_{tag}_recnrm_sha"] = np.array(sha(np.asarray(r["recnrm"], np.float64)))
        out[f"gs_{tag}_ijkgd_sha"] = np.array(sha(np.asarray(r["ijkgd"], np.int32)))
    # rnd00
    a, sa = R.rnd00(25.0, 1000)
    out["rnd00_first"] = np.asarray(a[:16], np.float64)
    out["rnd00_sha"] = np.array(sha(np.asarray(a, np.float64)))
    out["rnd00_seed_after"] = np.array(sa, np.float64)
    si8 = conftest.Sphere(1)
    # ekinetic2
    psi6 = initial_wavefunctions(si8, 6, 25.0)
    r2, e2, t2 = R.ekinetic2(psi6, occupations(6, 8), si8.dkg)
    out["ekinetic_ekib"] = np.asarray(e2)
    out["ekinetic_ekint"] = np.array(t2)
    # prpcg
    for i, ek in enumerate((0.7, 3.1)):
        out[f"prpcg_{i}"] = np.asarray(R.prpcg(ek, si8.dkg))[::SUB]
    # hckbpp_g / hckbpp_ib_g / enkbpp_g2
    nb = 5
    psi = initial_wavefunctions(si8, nb, 25.0)
    rh0 = initial_wavefunctions(si8, nb, 77.0)
    eq, lmx, vcl, cl = model_kb_projectors(si8, si8.cell)
    lmx = lmx.copy()
    lmx[1] = 1
    natom, lmmax = R.kb_setup(eq, lmx, vcl, cl, nb)
    r_ref, sl_ref = R.hckbpp_g(psi, rh0, natom, lmmax)
    out["hckbpp_rhcr"] = np.asarray(r_ref)[:, ::SUB]
    sl = np.array(sl_ref, copy=True)
    for ia in range(natom):
        sl[:, ia, lmx[ia]:] = 0.0
    out["hckbpp_slmir"] = sl
    occ = np.array([2.0, 2.0, 2.0, 0.0, 1.0])
    out["enkbpp"] = np.array(R.enkbpp_g2(nb, natom, occ, si8.cell.volume))
    nw = si8.nplw + 1
    out["kbnlpp_frc2"] = R.kbnlpp_frc2(psi, occ, si8.gx[:nw], si8.gy[:nw], si8.gz[:nw], natom)
    out["hckbpp_ib"] = np.asarray(R.hckbpp_ib_g(psi[2], rh0[2], natom))[::SUB]
    # schmidt2
    b, okb = R.schmidt2(initial_wavefunctions(si8, 12, 25.0))
    assert okb
    out["schmidt2"] = np.asarray(b)[:, ::SUB]
    # rsdecmp + rsinvcd
    rng = np.random.default_rng(3)
    x = rng.standard_normal((40, 25))
    out["icholesky"] = np.tril(R.icholesky(x.T @ x + 0.1 * np.eye(25)))
    # hmatrix2, untrycore
    p9 = initial_wavefunctions(si8, 9, 25.0)
    h9 = initial_wavefunctions(si8, 9, 5.0)
    hm, ehm = R.hmatrix2(p9, h9)
    out["hmatrix2"] = np.asarray(hm)
    out["hmatrix2_eig"] = np.asarray(ehm)
    u = np.linalg.qr(np.random.default_rng(1).standard_normal((9, 9)))[0]
    out["untrycore"] = np.asarray(R.untrycore(p9, u))[:, ::SUB]
    # RsEIGQR
    rng = np.random.default_rng(11)
    for n in (7, 40):
        xx = rng.standard_normal((n, n))
        w_ref, v_ref = R.rseigqr(xx + xx.T)
        out[f"rseigqr_{n}_w"] = np.asarray(w_ref)
        v = np.asarray(v_ref)
        out[f"rseigqr_{n}_v"] = v * np.sign(v[np.argmax(np.abs(v), axis=0), np.arange(n)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")
    # stscfn_pw and setnlc_g (two projector types) on Si8
    c = conftest.nlpp_case(si8)
    yc, ys = R.stscfn_pw(c["frac"], si8.cell.hcell, c["nd"], c["nd"], *c["g"], c["nplwcs"], c["kmax"])
    out["stscfn_ycos"] = np.asarray(yc)[:, ::SUB]
    out["stscfn_ysin"] = np.asarray(ys)[:, ::SUB]
    eqr, vclr = R.setnlc_g(c["tables"], c["nhk1"], c["nhk2"], c["natom"], yc, ys, c["recnrm"],
                           si8.cell.volume)
    out["setnlc_eqlmr"] = np.asarray(eqr)[:, :, ::SUB]
    out["setnlc_vclalp"] = np.asarray(vclr)
    # xclda (both Perdew-Zunger branches, core charge on part of the mesh) and INVERS
    from test_scf_glue import xc_inputs
    rho, core = xc_inputs()
    v, e, e4 = R.xclda(rho, core, 0.37)
    out["xclda_vexc"] = np.asarray(v)[::7]
    out["xclda_eeexc"] = np.asarray(e)[::7]
    out["xclda_e4"] = np.asarray(e4)
    rng = np.random.default_rng(21)
    for n in (2, 5, 9):
        x = rng.standard_normal((n, n))
        wi, ierr = R.invers(x @ x.T + 0.05 * np.eye(n))
        assert ierr == 0
        out[f"invers_{n}"] = np.asarray(wi)
    # flocal_g (two types, a permuted iatmpt, llking on the second type once)
    f = conftest.flocal_case(si8)
    for tag, llk in (("", [0, 0]), ("_lking", [0, 1])):
        out["flocal_g" + tag] = R.flocal_g(f["ntype"], f["nhk1"], f["nhk2"], f["iatmpt"], llk,
                                           f["natom"], si8.nga, si8.ngb, si8.ngc, si8.gx, si8.gy,
                                           si8.gz, f["rhgr"], f["svco"], f["kmax"], f["tables"],
                                           np.full((f["natom"], 3), 0.25))
    # frcrec (Ewald reciprocal part, VECTOR form)
    e = conftest.ewald_case(si8)
    fr, er, aer = R.frcrec(e["ntype"], e["nhk1"], e["nhk2"], e["zv"], e["nel"], e["natom"],
                           e["iatoit"], e["iatmpt"], e["ew"], e["tables"],
                           np.full((e["natom"], 3), 0.5), 1.25, np.full(e["natom"], -0.5))
    out["frcrec_fclm"] = fr
    out["frcrec_ecolmb"] = np.array(er)
    out["frcrec_atom_ecolmb"] = aer
    # roteig2 and set_coef_ASPC (wave-function extrapolation)
    cur, prev = conftest.roteig_case(si8)
    out["roteig_prev"] = np.asarray(R.roteig(cur, prev, 2))[:, :, ::SUB]
    for ih in (1, 2, 3, 4):
        out[f"coef_aspc_{ih}"] = R.coef_aspc(ih)
    # velocity-Verlet updates of the MD loop
    md = conftest.md_case()
    out["md_updtps"] = R.md_kick(0, md["x"], md["v"], md["species"], md["lfixion"])
    out["md_vkick"] = R.md_kick(1, md["v"], md["f"], md["species"], md["lfixion"])
    out["md_normalized_a"] = R.normalized_a(md["f"], md["species"], md["hcell"], md["acon"])
    out["md_updtps_pbc"] = R.updtps_pbc(md["x"])
    b, vol = R.rciprl(md["hcell"])
    out["md_rciprl_b"] = b
    out["md_rciprl_v"] = np.array(vol)
    # restart files written by the reference (svwfns both encodings, svacon):
    # their SHA-256 digests
    import tempfile
    g, eig, occ, ion = conftest.restart_case(si8)
    with tempfile.TemporaryDirectory() as d:
        for l8 in (True, False):
            f = os.path.join(d, "QM_eig")
            R.svwfns(f, g, eig, occ, l8)
            out[f"qm_eig_sha_{int(l8)}"] = np.array(sha(np.frombuffer(open(f, "rb").read(),
                                                                   np.uint8)))
        f = os.path.join(d, "QM_ion")
        R.svacon(f, ion["nhk1"], ion["nhk2"], ion["nstepcg"], ion["nstepmd"], ion["lclust"],
                 ion["ratm"], ion["ifmd"], ion["frc"], ion["prevr"], ion["hcell"])
        out["qm_ion_sha"] = np.array(sha(np.frombuffer(open(f, "rb").read(), np.uint8)))
    # flang's formatting of the text outputs' edit descriptors
    import shutil
    import subprocess
    fc = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "ff")
        subprocess.run([fc, os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                         "fortran_formats.f90"), "-o", exe], check=True)
        lines = subprocess.run([exe], stdout=subprocess.PIPE, text=True, check=True).stdout
    out["fortran_format_lines"] = np.array(lines.splitlines())
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
