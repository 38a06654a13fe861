"""Build qxmd_amd/libqxmd_amd.so (HIP, gfx950) in-tree.

Run as ``python -m qxmd_amd.build`` or through ``__graft_entry__.build()``.
Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled; the plane-wave set-up (gsphere.cpp) is host code compiled
without floating-point contraction so that its G-vector sort keys match the
reference bit for bit.  The Fortran host layer (qxmd_amd/fortran) is built
with flang into libqxmd_amd_f.so when flang is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqxmd_amd.so")
BLD = os.path.join(HERE, "_build")
ARCH = os.environ.get("QX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["pencil.hip", "pencil_fast.hip", "gemm.hip", "mfma_gemm.hip", "blaslt.hip", "dense.hip",
               "eigh.hip",
               "fullgrid.hip", "kpoint.hip", "cg.hip", "nlsetup.hip", "kb1.hip", "forces.hip", "flocal.hip", "ewald.hip",
               "aspc.hip", "scfglue.hip", "api.hip"]
HOST_SOURCES = ["gsphere.cpp", "restart.cpp", "mdstep.cpp"]
FORTRAN_SOURCES = ["qxmd_amd_c.f90", "qxmd_dropin.f90"]
FOUT = os.path.join(HERE, "libqxmd_amd_f.so")
ROOT = os.path.dirname(HERE)

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(BLD, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BLD, s.replace(".hip", ".o"))
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HOST_SOURCES:
        o = os.path.join(BLD, s.replace(".cpp", ".o"))
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-c",
                     os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)
    tmp = OUT + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lhipblaslt", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


def _flang():
    for c in (os.environ.get("FC_FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose: bool = False):
    """The Fortran host layer: ISO_C_BINDING interfaces + drop-in externals
    (libqxmd_amd_f.so next to libqxmd_amd.so) and the Fortran check driver
    tests/fortran/dropin_check.  Skipped when no flang is installed."""
    fc = _flang()
    if fc is None:
        if verbose:
            print("flang not found: Fortran layer skipped")
        return None
    fdir = os.path.join(HERE, "fortran")
    objs = []
    for src in FORTRAN_SOURCES:
        o = os.path.join(BLD, src.replace(".f90", ".o"))
        _run([fc, "-O2", "-fPIC", "-module-dir", BLD, "-c", os.path.join(fdir, src), "-o", o])
        objs.append(o)
    tmp = FOUT + ".tmp"
    _run([fc, "-shared", "-o", tmp, *objs, "-L" + HERE, "-lqxmd_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FOUT)
    build_module_dropins(fc, verbose)
    drv_src = os.path.join(ROOT, "tests", "fortran", "dropin_check.f90")
    if os.path.exists(drv_src):
        drv = os.path.join(ROOT, "tests", "fortran", "dropin_check")
        _run([fc, "-O2", "-I" + BLD, drv_src, "-o", drv + ".tmp", "-L" + HERE, "-lqxmd_amd_f",
              "-lqxmd_amd", "-Wl,-rpath,$ORIGIN/../../qxmd_amd"])
        os.replace(drv + ".tmp", drv)
    return FOUT


def build_module_dropins(fc, verbose=False):
    """qxmd_dropin_qxmd.f90 (schmidt, pwpchg and eigen by name) reads QXMD's
    own modules, so it is compiled inside a QXMD build.  Where the
    reference's flang build of the checker (oracle/_ref) is present, it is
    compiled here against that build's modules and linked into the check
    driver tests/fortran/dropin_modules_check (the reference's module data
    comes from oracle/_ref/libqxref.so; the executable's schmidt / pwpchg /
    eigen and libqxmd_amd_f.so's softchg / chgr2g take precedence over the
    reference's).  Skipped elsewhere."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    refmod = os.path.join(ref, "build")
    if not (os.path.exists(os.path.join(ref, "libqxref.so"))
            and os.path.exists(os.path.join(refmod, "ifcholesky.mod"))):
        if verbose:
            print("oracle/_ref not built: module drop-in check skipped")
        return None
    mdir = os.path.join(BLD, "qxmd_modules")
    os.makedirs(mdir, exist_ok=True)
    obj = os.path.join(mdir, "qxmd_dropin_qxmd.o")
    _run([fc, "-O2", "-fPIC", "-I" + refmod, "-I" + BLD, "-module-dir", mdir, "-c",
          os.path.join(HERE, "fortran", "qxmd_dropin_qxmd.f90"), "-o", obj])
    src = os.path.join(ROOT, "tests", "fortran", "dropin_modules_check.f90")
    drv = os.path.join(ROOT, "tests", "fortran", "dropin_modules_check")
    dobj = os.path.join(mdir, "dropin_modules_check.o")
    _run([fc, "-O2", "-I" + refmod, "-I" + BLD, "-module-dir", mdir, "-c", src, "-o", dobj])
    _run([fc, dobj, obj, "-o", drv + ".tmp", "-L" + HERE, "-lqxmd_amd_f", "-lqxmd_amd", "-L" + ref,
          "-lqxref", "-Wl,--allow-shlib-undefined",
          "-Wl,-rpath,$ORIGIN/../../qxmd_amd:$ORIGIN/../../oracle/_ref"])
    os.replace(drv + ".tmp", drv)
    return drv


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
