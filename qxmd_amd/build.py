"""Build qxmd_amd/libqxmd_amd.so (HIP, gfx950) in-tree.

Run as ``python -m qxmd_amd.build`` or through ``__graft_entry__.build()``.
Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled; the plane-wave set-up (gsphere.cpp) is host code compiled
without floating-point contraction so that its G-vector sort keys match the
reference bit for bit.
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

HIP_SOURCES = ["pencil.hip", "pencil_fast.hip", "gemm.hip", "dense.hip", "eigh.hip", "fullgrid.hip", "cg.hip", "api.hip"]
HOST_SOURCES = ["gsphere.cpp"]

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
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
