"""The reference's per-step text outputs, written and read in its formats:
qm_eng.d (energies of each SCF iteration, Sources/pwlda.f90:2592-2598,
format (2i6,2es18.10,20es14.6)), qm_eig.d (eigenvalues and occupations,
Sources/fermi.f90:1924-1937, '(10i7)' then '(i6,es13.5,f6.3)' per band, f7.3
for negative occupations) and qm_fer.d (Fermi energy, '(2i7,2es13.5)').
The writers reproduce the Fortran edit descriptors so that files from this
library and from the reference compare line by line."""
from __future__ import annotations

import numpy as np

# the energy columns of qm_eng.d after nstep and the SCF iteration count
ENG_COLUMNS = ("sume", "etot", "ekin", "eext", "ehar", "eexc", "ecor", "ecnl", "entrpy",
               "eonsite", "exhfsr", "eplusu", "eefield", "soenergy")


def fortran_es(x, w, d):
    """ESw.d: one digit before the point, d after, two-digit exponent."""
    return f"{float(x):{w}.{d}E}"


def fortran_f(x, w, d):
    return f"{float(x):{w}.{d}f}"


def eng_line(nstep, jgcycl, values):
    """One qm_eng.d line: values in ENG_COLUMNS order (ecor is written as
    ecor - ecnl by the reference; pass the column as written)."""
    v = list(values)
    return (f"{int(nstep):6d}{int(jgcycl):6d}" + "".join(fortran_es(x, 18, 10) for x in v[:2])
            + "".join(fortran_es(x, 14, 6) for x in v[2:]))


def eig_block(nstep, jgcycl, eig, occ, order=None):
    """The qm_eig.d lines of one step (Gamma point): the band index written is
    the reference's norder (1-based) when given, else 1..nband."""
    nb = len(eig)
    order = np.arange(1, nb + 1) if order is None else np.asarray(order)
    lines = [f"{int(nstep):7d}{int(jgcycl):7d}{nb:7d}"]
    for jb in order:
        o = float(occ[jb - 1])
        lines.append(f"{int(jb):6d}" + fortran_es(eig[jb - 1], 13, 5)
                     + (fortran_f(o, 6, 3) if o >= 0.0 else fortran_f(o, 7, 3)))
    return lines


def fer_line(nstep, jgcycl, fermie):
    return f"{int(nstep):7d}{int(jgcycl):7d}" + fortran_es(fermie, 13, 5)


def read_qm_eng(path):
    """Rows of qm_eng.d: (nstep, jgcycl, dict of the energy columns)."""
    rows = []
    for line in open(path):
        if not line.strip() or line.lstrip().startswith("#"):
            continue
        f = line.split()
        vals = [float(x) for x in f[2:]]
        rows.append((int(f[0]), int(f[1]), dict(zip(ENG_COLUMNS, vals))))
    return rows


def read_qm_eig(path):
    """Steps of qm_eig.d: (nstep, jgcycl, band index, eig, occ) arrays per step."""
    steps = []
    lines = [ln for ln in open(path) if ln.strip() and not ln.lstrip().startswith("#")]
    i = 0
    while i < len(lines):
        h = lines[i].split()
        nstep, jg, nb = int(h[0]), int(h[1]), int(h[2])
        rows = [lines[i + 1 + k].split() for k in range(nb)]
        idx = np.array([int(r[0]) for r in rows])
        steps.append((nstep, jg, idx, np.array([float(r[1]) for r in rows]),
                      np.array([float(r[2]) for r in rows])))
        i += 1 + nb
    return steps


def read_qm_fer(path):
    """Rows of qm_fer.d: (nstep, jgcycl, fermie)."""
    out = []
    for line in open(path):
        if not line.strip() or line.lstrip().startswith("#"):
            continue
        f = line.split()
        out.append((int(f[0]), int(f[1]), float(f[2])))
    return out
