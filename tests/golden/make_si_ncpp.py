"""Write tests/golden/si_ncpp_lda.npz: the silicon norm-conserving
pseudopotential (the reference's LIB/NCPP/SI.PBE files with the functional
marker of SI_A.rvs set to -1, so that LDA accepts them; Sources/ppkb.f90:
1181-1189) as the reference's own set-up routines turn it into tables
(read_pp, rdtmpp, lotmpp, lotmpp_g, nlkbpp_g; driven through
oracle/_ref/libqxref.so by oracle.ref.ncpp).

Per cell (diamond Si, a0 = 10.26 bohr, n x n x n; "tiny" is the 8-atom cell
at 6 / 24 Ry, the others at 20 / 80 Ry) the fixture holds what depends on
the cell's sphere only through its largest |G|:
    <cell>_tbqlm, _tbqlma (lnref, 551), _dlqlm     KB radial table (nlkbpp_g)
    <cell>_rhocr, _rhocra (1001)                   local radial table (lotmpp_g)
and once: zv, lmx, lmmax, lnref, nqlmmx, clalp, itola, itolm, itolmf, litoll
(entries past lmx zeroed: the reference leaves them unset and never reads
them).  For the 8-atom cells it also holds the per-G outputs that the host
restatements in qxmd_amd/ncpp.py must reproduce: qlmr / qlmi (angular
factors, nlkbpp_g) and svco (local form factors, lotmpp_g).

    python tests/golden/make_si_ncpp.py      (needs oracle/_ref built)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle.ref as R  # noqa: E402

PPDIR = "/root/reference/LIB/NCPP/SI.PBE"
CELLS = {"tiny": (1, 6.0, 24.0), "si8": (1, 20.0, 80.0), "si64": (2, 20.0, 80.0),
         "si216": (3, 20.0, 80.0), "si512": (4, 20.0, 80.0)}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "si_ncpp_lda.npz")


def main():
    if not R.available():
        raise SystemExit("oracle/_ref is not built (make -f oracle/ref.mk)")
    from conftest import Sphere
    out = {}
    common = None
    for tag, (n, ecut, ecutdens) in CELLS.items():
        s = Sphere(n, ecut=ecut, ecutdens=ecutdens)
        nw, n5 = s.nplw + 1, s.nplw5 + 1
        r = R.ncpp(14.0, "si", PPDIR, 1, ecutdens, s.gx[:nw], s.gy[:nw], s.gz[:nw],
                   s.recnrm[:nw], s.recnrm[:n5])
        lmx = r["lmx"]
        for k in ("clalp", "itola", "itolm", "itolmf", "litoll"):
            r[k] = np.asarray(r[k]).copy()
            r[k][lmx:] = 0
        c = {k: r[k] for k in ("zv", "lmx", "lmmax", "lnref", "nqlmmx", "lmax", "lclno", "clalp",
                               "itola", "itolm", "itolmf", "litoll")}
        if common is None:
            common = c
        else:
            for k, v in c.items():
                assert np.array_equal(np.asarray(v), np.asarray(common[k])), (tag, k)
        out[f"{tag}_tbqlm"] = r["tbqlm"]
        out[f"{tag}_tbqlma"] = r["tbqlma"]
        out[f"{tag}_dlqlm"] = np.array(r["dlqlm"])
        out[f"{tag}_rhocr"] = r["rhocr"]
        out[f"{tag}_rhocra"] = r["rhocra"]
        if n == 1:
            out[f"{tag}_qlmr"] = r["qlmr"]
            out[f"{tag}_qlmi"] = r["qlmi"]
            out[f"{tag}_svco"] = r["svco"]
        print(tag, "nplw", s.nplw, "nplw5", s.nplw5, "dlqlm", r["dlqlm"], "svco0",
              r["svco"][0])
    for k, v in common.items():
        out[k] = np.asarray(v)
    out["provenance"] = np.array(
        "reference QXMD read_pp/rdtmpp/lotmpp/lotmpp_g/nlkbpp_g (flang build, oracle/ref.mk) on "
        "LIB/NCPP/SI.PBE with the SI_A.rvs functional marker -4 -> -1 (LDA, jgga = 1)")
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
