"""The reference's per-step text outputs, written and read in its formats:
qm_eng.d (energies of each SCF iteration, Sources/pwlda.f90:2592-2598,
format (2i6,2es18.10,20es14.6)), qm_eig.d (eigenvalues and occupations,
Sources/fermi.f90:1924-1937, '(10i7)' then '(i6,es13.5,f6.3)' per band, f7.3
for negative occupations), qm_fer.d (Fermi energy, '(2i7,2es13.5)'), and
qm_frc.d / qm_ion.d (forces and scaled coordinates of each MD step,
Sources/datadump.f90:235-266: '(100i7)' nstep, ntype, nhk(1:ntype), then the
scale '(es14.7)', then the values over the scale '(9f8.5)').
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


# ---- qm_frc.d / qm_ion.d ---------------------------------------------------
FRC_HEADER = "#  Atomic forces in [a.u.]       "
ION_HEADER = "#  Atomic scaled coordinates     "


def _i7_lines(vals, per=100):
    vals = [int(v) for v in vals]
    return ["".join(f"{v:7d}" for v in vals[k:k + per]) for k in range(0, len(vals), per)]


def _f85_lines(vals):
    vals = [float(v) for v in vals]
    return ["".join(fortran_f(v, 8, 5) for v in vals[k:k + 9]) for k in range(0, len(vals), 9)]


def frc_block(nstep, nhk, frc):
    """The qm_frc.d lines of one MD step: frc (natom, 3) in a.u., nhk the atom
    count of each species (Sources/datadump.f90:246-266)."""
    f = np.asarray(frc, np.float64).reshape(-1, 3)
    atconfx = float(np.abs(f).max()) / 9.999 if f.size else 0.0
    lines = _i7_lines([nstep, len(nhk), *nhk])
    lines.append(fortran_es(atconfx, 14, 7))
    if abs(atconfx) < 1e-30:  # the reference's guard against dividing by zero
        atconfx = 1.0
    return lines + _f85_lines(x / atconfx for x in f.reshape(-1))


def ion_block(nstep, nhk, vatm):
    """The qm_ion.d lines of one MD step: scaled coordinates vatm (natom, 3)
    written over 0.1 and capped at 9.99999 (Sources/datadump.f90:235-241)."""
    v = np.asarray(vatm, np.float64).reshape(-1)
    lines = _i7_lines([nstep, len(nhk), *nhk])
    lines.append(fortran_es(0.1, 14, 7))
    return lines + _f85_lines(min(x / 0.1, 9.99999) for x in v)


def _read_blocks(path):
    """Steps of a qm_frc.d / qm_ion.d file: (nstep, nhk, scale, values (natom,
    3) as written).  The (9f8.5) fields can touch, so they are cut by width."""
    lines = [ln.rstrip("\n") for ln in open(path)
             if ln.strip() and not ln.lstrip().startswith("#")]
    steps, i = [], 0
    while i < len(lines):
        h = [int(lines[i][k:k + 7]) for k in range(0, len(lines[i].rstrip()), 7)]
        nstep, ntype = h[0], h[1]
        while len(h) < 2 + ntype:  # more than 100 integers continue on the next line
            i += 1
            h += [int(lines[i][k:k + 7]) for k in range(0, len(lines[i].rstrip()), 7)]
        nhk = np.array(h[2:2 + ntype])
        scale = float(lines[i + 1])
        nval = 3 * int(nhk.sum())
        vals = []
        i += 2
        while len(vals) < nval:
            ln = lines[i]
            vals += [float(ln[k:k + 8]) for k in range(0, len(ln.rstrip()), 8)]
            i += 1
        steps.append((nstep, nhk, scale, np.array(vals[:nval]).reshape(-1, 3)))
    return steps


def read_qm_frc(path):
    """Steps of qm_frc.d: (nstep, nhk, forces (natom, 3) in a.u.), the forces
    as written (5 decimals of force / scale) times the scale."""
    return [(n, nhk, vals * scale) for n, nhk, scale, vals in _read_blocks(path)]


def read_qm_ion(path):
    """Steps of qm_ion.d: (nstep, nhk, scaled coordinates (natom, 3)); values
    written at the 9.99999 cap read back as 0.999999."""
    return [(n, nhk, vals * scale) for n, nhk, scale, vals in _read_blocks(path)]
