#!/usr/bin/env python
"""Benchmark: Gamma-point H|psi> band-applies/s and SCF band-step time on the
216-atom Si cell (3x3x3 diamond, a = 30.78 bohr, 20/80 Ry, 90^3 grid, 21909
half-sphere plane waves, 540 bands, 864 KB projectors), mixed precision vs FP64.

A step is one batched H|psi> over all 540 bands (540 band-applies: kinetic,
sphere-pruned FFT pair with V(r), KB projection and back-projection) with the
inputs resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--ncell 3] [--no-scf] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3
BF16_PEAK_TFS = 2500.0


NCPP = os.path.join(ROOT, "tests", "golden", "si_ncpp_lda.npz")
CELL_TAG = {1: "si8", 2: "si64", 3: "si216", 4: "si512"}


def ewald_list(gs, hcell, gamma=0.35):
    """Ewald G list from the reference's clmrcp generator (its order): the
    half space of reciprocal vectors with (pi |G| / gamma)^2 <= 37, in the
    units clmrcp uses (no 2 pi)."""
    from qxmd_amd import pw
    ew = pw.clmrcp(np.linalg.inv(hcell).T, 37.0 * gamma ** 2 / np.pi ** 2)
    n = len(ew["nga"])
    return dict(nga=ew["nga"], ngb=ew["ngb"], ngc=ew["ngc"], gx=ew["gx"], gy=ew["gy"],
                gz=ew["gz"], recnrm=ew["recnrm"], recnrmex=np.zeros(n + 1), gamma=gamma,
                volume=gs.volume)


class Workload:
    """The silicon cell of BASELINE's configurations: n x n x n diamond cells
    (a0 = 10.26 bohr) with seeded displacements, 20 / 80 Ry, the reference's
    Si LDA norm-conserving pseudopotential (tables from its own set-up
    routines), seeded damped random starting bands."""

    def __init__(self, ncell, nband):
        from qxmd_amd import ncpp, pw
        from qxmd_amd.system import diamond_si, initial_wavefunctions, occupations
        self.cell = diamond_si(ncell)
        self.gs = pw.GSphere(self.cell.hcell)
        self.natom = len(self.cell.tau)
        self.nband = nband
        self.tables, self.svco, self.zv = ncpp.silicon_cell_inputs(self.gs, np.load(NCPP),
                                                                   CELL_TAG[ncell])
        self.frac = np.linalg.solve(self.cell.hcell, self.cell.tau.T).T
        self.psi0 = initial_wavefunctions(self.gs, nband)
        self.occ = occupations(nband, 4 * self.natom)
        self.ew = ewald_list(self.gs, self.cell.hcell)

    def scf(self, ctx, x=None):
        """A fresh SCF driver on ctx (new Pulay history) at positions x."""
        from qxmd_amd.scf import Scf
        sc = Scf(ctx, self.natom, self.tables, self.svco, self.zv)
        sc.set_ewald(self.ew)
        sc.set_positions(self.frac if x is None else x)
        return sc


def time_host(ctx, fn, reps):
    fn()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    ctx.sync()
    return round((time.perf_counter() - t0) / reps * 1e3, 4)


def md_parts(ctx, w, sc, reps=3):
    """Per-MD-step pieces on the device with the SCF's own state: the refresh
    at new positions (stscfn_pw, setnlc_g's KB table, vlocal_g), the SCF glue
    between band steps (outchg, Pulay mixing with the history held,
    potential), the KB / local / reciprocal-Ewald forces."""
    from qxmd_amd import pw
    nb, occ = w.nband, w.occ
    out = {"refresh_ms": time_host(ctx, lambda: sc.set_positions(w.frac), reps)}
    ctx.put(0, w.psi_orth)
    sc.start(None, occ)
    for it in range(4):  # fill part of the Pulay history
        ctx.scf_band_step(nb, occ, w.gs.volume, prec=pw.MIXED, batched=True, want_rho=False)
        ctx.outchg()
        ctx.pwpmxchgg(sc.aslh, sc.bslh, lreset=(it == 0))
        ctx.scf_potential()

    def glue():
        ctx.outchg()
        ctx.pwpmxchgg(sc.aslh, sc.bslh, lmixhold=True)
        ctx.scf_potential()

    out["scf_glue_ms"] = time_host(ctx, glue, reps)
    out["pwpchg_ms"] = time_host(ctx, lambda: ctx.pwpchg_slot(0, nb, occ), reps)
    out["kb_forces_ms"] = time_host(ctx, lambda: ctx.kb_forces(0, nb, occ), reps)
    one, nat = sc.nhk1, sc.nhk2
    out["flocal_ms"] = time_host(
        ctx, lambda: ctx.flocal_g_dev(sc.ntype, one, nat, sc.ions, sc._zero, w.natom), reps)
    it = np.ones(w.natom, np.int32)
    out["ewald_frcrec_ms"] = time_host(
        ctx, lambda: ctx.frcrec(sc.ntype, one, nat, sc.zv, 4 * w.natom, w.natom, it, sc.ions),
        reps)
    out["ewald_g"] = len(w.ew["nga"])
    return out


def extrapolation(ctx, psi, nband, reps=3):
    """Time sbalinaspc (ihest = 2, three previous sets): the bands of each
    step are psi moved slightly out of its span and orthonormalised again
    (untimed, as the SCF leaves them), so the overlap, its polar factor, the
    rotation of the history and Gram-Schmidt all do real work."""
    rng = np.random.default_rng(3)
    ctx.aspc_init(3, 2)
    noise = 1e-3 * rng.standard_normal(psi.shape)
    times = []
    for step in range(1, 4 + reps):
        ctx.put(0, psi + step * noise)
        assert ctx.schmidt(0, nband)
        ctx.sync()
        t0 = time.perf_counter()
        ia, ok = ctx.sbalinaspc(0, nband, step)
        dt = time.perf_counter() - t0
        if step > 3:
            times.append(dt)
    return {"step_ms": round(1e3 * float(np.mean(times)), 3), "iaspc": ia, "lcholesky": ok}


def scf_converge(ctx, w, prec, batched, refine=None, tol=1e-6, max_scf=80):
    """SCF from the shared start (orthonormal seeded bands, their density) to
    a change of the Harris-Foulkes energy below tol Ha/atom; then the final
    Rayleigh-Ritz and the forces.  Timed from the first band step to the
    forces."""
    sc = w.scf(ctx)
    ctx.put(0, w.psi_orth)
    sc.start(None, w.occ)
    ctx.sync()
    t0 = time.perf_counter()
    it, conv, hist = sc.iterate(w.occ, tol_ha_atom=tol, max_scf=max_scf, prec=prec,
                                batched=batched, refine_ha_atom=refine)
    eig, e = sc.finish(w.occ)
    frc = sc.forces(w.occ)
    ctx.sync()
    dt = time.perf_counter() - t0
    return {"ms": round(dt * 1e3, 2), "iterations": it, "converged": conv,
            "fp64_refinement_steps": sc.fp64_steps, "energy_ry": e, "eig": eig, "forces": frc}


def scf_errors(r, ref, natom):
    """Errors of run r against ref (Ha units): energy per atom, eigenvalues of
    the occupied bands, forces."""
    nocc = 2 * natom
    return {"energy_ha_per_atom": float(abs(r["energy_ry"] - ref["energy_ry"]) / 2 / natom),
            "max_occ_eig_ha": float(np.abs(r["eig"][:nocc] - ref["eig"][:nocc]).max() / 2),
            "max_force_ha_bohr": float(np.abs(r["forces"] - ref["forces"]).max() / 2)}


def md_run(ctx, w, nsteps=3, max_scf=40, tol_ha_atom=1e-6, dt=41.341, temperature=300.0,
           seed=7):
    """Converged MD steps at constant energy (Si masses, dt in a.u.: 1 fs),
    from seeded Maxwell-Boltzmann velocities at `temperature` (K, centre of
    mass at rest), each SCF in mixed precision with the final FP64 refinement
    sweep, device-resident except the O(natom) Verlet updates.  Per step:
    half kick and position update (host); structure factors, KB table and
    local potential at the new positions (device); sbalinaspc; SCF
    iterations until the Harris-Foulkes energy changes by less than
    tol_ha_atom per atom; KB, local and reciprocal-space Ewald forces (the
    Ewald real-space sum is not part of the path); second half kick.  Step 0
    converges from the starting bands.  After the last step an FP64 batched
    SCF is continued at the same positions to 1e-8 Ha/atom, and the last
    step's forces and energy are compared with its."""
    from qxmd_amd import md, pw
    natom, nb, occ = w.natom, w.nband, w.occ
    hcell = w.cell.hcell
    species = np.ones(natom, np.int32)
    lfix = np.zeros(1, bool)
    mass = 28.0855 * 1822.888486
    acon = md.acon([mass], dt)
    x = w.frac.copy()
    # velocities: Cartesian components N(0, kT/m) in a.u., the mean removed,
    # as displacements per step in cell units (the Verlet routines' v)
    rng = np.random.default_rng(seed)
    vc = rng.normal(0.0, np.sqrt(3.166811563e-6 * temperature / mass), (natom, 3))
    vc -= vc.mean(axis=0)
    v = (vc * dt) @ np.linalg.inv(hcell).T
    sc = w.scf(ctx, x)
    ctx.put(0, w.psi_orth)
    ctx.aspc_init(3, 2)
    sc.start(None, occ)
    ctx.sync()
    t0 = time.perf_counter()
    n0, conv0, _ = sc.iterate(occ, tol_ha_atom, max_scf, pw.MIXED, True,
                              refine_ha_atom=tol_ha_atom)
    a = md.normalized_a(sc.forces(occ), species, hcell, acon)
    v = md.vkick(v, a, species, lfix)
    ctx.sync()
    first = {"scf_iterations": n0, "converged": conv0,
             "ms": round((time.perf_counter() - t0) * 1e3, 2)}
    steps, iters, convs = [], [], []
    f = e = None
    for k in range(1, nsteps + 1):
        ctx.sync()
        t0 = time.perf_counter()
        v = md.vkick(v, a, species, lfix)
        x = md.updtps_pbc(md.updtps(x, v, species, lfix))
        sc.set_positions(x)
        ctx.sbalinaspc(0, nb, k)
        n, conv, hist = sc.iterate(occ, tol_ha_atom, max_scf, pw.MIXED, True,
                                   refine_ha_atom=tol_ha_atom)
        f = sc.forces(occ)
        e = hist[-1]
        a = md.normalized_a(f, species, hcell, acon)
        v = md.vkick(v, a, species, lfix)
        ctx.sync()
        steps.append((time.perf_counter() - t0) * 1e3)
        iters.append(n)
        convs.append(conv)
    out = {"step_ms": round(float(np.mean(steps)), 2), "steps": nsteps, "dt_fs": 1.0,
           "temperature_k": temperature, "scf_iterations": iters, "converged": all(convs),
           "tol_ha_atom": tol_ha_atom, "scf_prec": "mixed + FP64 refinement",
           "first_step": first}
    if f is not None:
        n_ref, conv_ref, hist_ref = sc.iterate(occ, 1e-8, max_scf, pw.FP64, True)
        f_ref = sc.forces(occ)
        out["last_step_vs_fp64"] = {
            "reference": "FP64 batched SCF continued at the last step's positions to 1e-8 Ha/atom",
            "reference_scf_iterations": n_ref, "reference_converged": conv_ref,
            "max_force_err_ha_bohr": float(np.abs(f - f_ref).max()) / 2.0,
            "energy_err_ha_atom": float(abs(e - hist_ref[-1])) / 2.0 / natom}
    return out


def fft_stage_lower_bytes(gs, ncol, nb):
    """Lower bound of the HBM bytes one launch of the fused mixed FFT stage
    (k_zinv_f + k_plane_f + k_zfwd_f over nb bands) must move in this design:
    per band the sphere psi in and H psi out (FP64, 8 nplwex bytes each), the
    FP32 copy of psi handed to the KB products and their FP32 result read
    back (4 nplwex each), and the z-pencil intermediates T1 and T2, each
    written once and read once (ncol columns x n points, FP32 complex, per
    band pair: 16 ncol n bytes per band); V(r) once per launch (FP32).  The
    kinetic term's second read of psi and V(r) re-reads beyond the first are
    not counted: the design could avoid them."""
    n1, n2, n3 = gs.nd
    per_band = 24 * gs.nplwex + 16 * ncol * n3
    return nb * per_band + 4 * n1 * n2 * n3, per_band


def time_region(ctx, fn, reps):
    ctx.sync()
    ctx.record(0)
    for _ in range(reps):
        fn()
    ctx.record(1)
    ctx.sync()
    return ctx.elapsed_ms(0, 1) / reps


def cpu_worker(ncell, budget_s):
    """Oracle H|psi> (FP64) on the host over the workload's bands, for as many
    as fit the budget (about budget_s of CPU work).  The potential and the KB
    table are the model ones of qxmd_amd.system (same grid, sphere and
    projector count, so the same operation count)."""
    import oracle as O
    from qxmd_amd import pw
    from qxmd_amd.system import (diamond_si, initial_wavefunctions, model_kb_projectors,
                                 model_local_potential)
    nband = int(2.5 * 8 * ncell ** 3)
    cell = diamond_si(ncell)
    gs = pw.GSphere(cell.hcell)
    v = model_local_potential(gs, cell)
    eq, lmx, vcl, _ = model_kb_projectors(gs, cell)
    psi = initial_wavefunctions(gs, nband)
    H = O.Hamiltonian(gs.nplw, gs.nd, gs.ijkgd(), v, gs.dkgnrm(), eq, lmx, vcl)
    H.hpsi(psi[:1])  # warm the twiddle tables
    t0 = time.perf_counter()
    n = 0
    while n + 2 <= nband:
        H.hpsi(psi[n:n + 2])
        n += 2
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    print(json.dumps({"bands": n, "seconds": dt}))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ncell, threads, budget_s=12.0):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                        "--ncell", str(ncell), "--cpu-budget", str(budget_s)],
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=600)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def kpoint_bench(steps, nb=160):
    """The 64-atom cell at the 8 k-points {0, 1/2}^3 (complex bands, each
    k-point its own sphere), nb bands each, FP64: H_k|psi> of all 8 x nb
    (k-point, band) blocks with one launch per stage, against the per-band
    launch loop.  Local potential: a smooth synthetic V(r); projectors: the
    reference's Si tables at k+G."""
    from qxmd_amd import kpoints, ncpp, pw
    from qxmd_amd.system import diamond_si
    cell = diamond_si(2)
    gs = pw.GSphere(cell.hcell)
    ctx = pw.PlaneWave(gs, 2)
    n1, n2, n3 = gs.nd
    x = np.arange(n1) / n1
    V = -0.3 + 0.05 * np.cos(2 * np.pi * (x[None, None, :] + 2 * x[None, :, None]
                                           + 3 * x[:, None, None]))
    ctx.set_glocal(np.ascontiguousarray(V.ravel()))
    b = 2 * np.pi * np.linalg.inv(cell.hcell).T
    kf = kpoints.gamma_folding_kpoints(2)
    npw = ctx.kpoints_set(kf, b, 20.0, nb)
    t, _, _ = ncpp.silicon_cell_inputs(gs, np.load(NCPP), "si512")
    rng = np.random.default_rng(7)
    for ik in range(len(kf)):
        hkl, kin = ctx.kpoints_sphere(ik)
        beta, D = kpoints.kb_projectors(t, cell.tau, np.zeros(cell.natom, int), b, kf[ik], hkl,
                                        cell.volume)
        ctx.kset_kb(ik, beta, D)
        c = (rng.standard_normal((nb, len(kin))) + 1j * rng.standard_normal((nb, len(kin))))
        ctx.kput(0, ik, c / (1.0 + kin)[None, :])
    ctx.khpsi(0, 1)  # first use (GEMM grid tuning)
    ctx.khpsi(0, 2, per_band=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.khpsi(0, 1)
    ctx.sync()
    ms_b = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    ctx.khpsi(0, 2, per_band=True)
    ctx.sync()
    ms_l = (time.perf_counter() - t0) * 1e3
    diff = max(float(np.abs(ctx.kget(1, ik) - ctx.kget(2, ik)).max()) for ik in range(len(kf)))
    # band-solver steps (block Rayleigh-Ritz on [psi, W] for all k-points):
    # one untimed start (buffers, first-use tuning), then 5 timed steps
    ctx.ksolve(0, 1, max_iter=1, tol=0.0)
    ctx.sync()
    t0 = time.perf_counter()
    _, nsteps = ctx.ksolve(0, 1, max_iter=5, tol=0.0)
    ctx.sync()
    ms_solve = (time.perf_counter() - t0) * 1e3
    ctx.close()
    nkb = len(kf) * nb
    return {"workload": f"Si64, 8 k-points {{0,1/2}}^3, {nb} bands each, FP64, grid {n1}^3",
            "npw_per_k": [int(v) for v in npw], "npj": int(t.lmx[0]) * cell.natom,
            "batched_ms": round(ms_b, 3), "batched_band_applies_per_s": round(nkb / ms_b * 1e3, 1),
            "per_band_loop_ms": round(ms_l, 2),
            "per_band_loop_band_applies_per_s": round(nkb / ms_l * 1e3, 1),
            "speedup_batched_vs_loop": round(ms_l / ms_b, 2),
            "max_abs_diff_batched_vs_loop": diff,
            "band_solver_ms_per_step": round(ms_solve / (nsteps + 1), 2),
            "band_solver_note": "one step = H_k of all 8 x nb residual directions + the two "
                                "stacked Rayleigh-Ritz products per k-point + one batched "
                                "rocSOLVER zhegvd (2 nb) + the two rotations; timed over the "
                                "start Rayleigh-Ritz and 5 steps"}


# ---- hardware counters of the FFT stage, measured in this run --------------
FFT_REGEX = "k_zinv_f|k_plane_f|k_zfwd_f|k_slot_copy"
PMC_PASSES = {
    "fetch": (["FETCH_SIZE"], FFT_REGEX),
    "write": (["WRITE_SIZE"], FFT_REGEX),
    "sq": (["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
            "SQ_ACTIVE_INST_LDS", "SQ_BUSY_CYCLES"], FFT_REGEX),
    # MFMA work and busy cycles of our GEMM kernels (MOPS counters count
    # units of 512 flops)
    # MFMA work and busy share of our GEMM kernels: rocprofv3's derived
    # MfmaUtil = sum SQ_VALU_MFMA_BUSY_CYCLES / (max GRBM_GUI_ACTIVE x SIMDs) x 100
    # and MfmaFlops* = SQ_INSTS_VALU_MFMA_MOPS_* x 512
    "mfma": (["MfmaUtil", "MfmaFlopsBF16", "MfmaFlopsF64"], "k_mg_bf3|k_mg_ps3|k_mg_f64"),
}
PMC_KERNELS = ("k_zinv_f<float", "k_plane_f<float", "k_zfwd_f<float")
MG_KERNELS = {"k_mg_ps3<true": "bf3_tn", "k_mg_ps3<false": "bf3_nn",
              "k_mg_bf3<float, float, true>": "bf3_tn", "k_mg_bf3<float, float, false>": "bf3_nn",
              "k_mg_f64<true": "f64_tn", "k_mg_f64<false": "f64_nn"}


def pmc_worker(ncell, nband, steps):
    """Runs under rocprofv3 --pmc: a device copy of all bands (slot 0 -> slot
    5: known bytes each way, the counters' calibration), then `steps` mixed
    H|psi> launches of the workload; the counters are read per dispatch."""
    from qxmd_amd import pw
    w = Workload(ncell, nband)
    ctx = pw.PlaneWave(w.gs, nband)
    sc = w.scf(ctx)
    ctx.put(0, w.psi0)
    sc.start(None, w.occ)
    ctx.slot_copy(0, 5, nband)  # nband * nplwex * 8 bytes read and written
    ctx.sync()
    for _ in range(steps):
        ctx.hpsi(0, 1, nband, pw.MIXED)
    ctx.sync()
    for _ in range(steps):  # the FP64 KB products (k_mg_f64) for the MFMA counters
        ctx.hpsi(0, 1, nband, pw.FP64)
    ctx.sync()
    ctx.close()
    print("pmc worker done", flush=True)


def _pmc_rows(outdir):
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            rows.extend(csv.DictReader(f))
    return rows


def _pmc_last(rows, keep):
    """Counter values of the last `keep` dispatches of each kernel of the
    stage (the worker's timed launches; earlier dispatches are the tuners'
    trial launches at other shapes), and of the last slot copy."""
    per = {}
    for row in rows:
        k = row.get("Kernel_Name", "")
        kern = next((p.split("<")[0] for p in PMC_KERNELS if p in k), None)
        if kern is None and "k_slot_copy" in k:
            kern = "copy"
        if kern is None:
            kern = next((v for n, v in MG_KERNELS.items() if n in k), None)
        if kern is None:
            continue
        did = int(row.get("Dispatch_Id") or row.get("Correlation_Id") or 0)
        per.setdefault((kern, row["Counter_Name"]), {})
        d = per[(kern, row["Counter_Name"])]
        d[did] = d.get(did, 0.0) + float(row["Counter_Value"])
    out = {}
    for (kern, cn), d in per.items():
        ids = sorted(d)[-(1 if kern == "copy" else keep):]
        out[(kern, cn)] = [d[i] for i in ids]
    return out


def _pmc_pass(cmd, outdir, limit_s):
    """Run one counter pass as a child process group.  rocprofv3 writes its
    CSV at the tool's finalisation, after the worker's last line; a pass whose
    CSV has been written and stopped growing is ended there (the profiler's
    own exit can stall in its signal handlers after that).  Returns (rows,
    worker finished, seconds, return code)."""
    import signal
    logp = outdir + ".log"
    os.makedirs(os.path.dirname(logp), exist_ok=True)
    t0 = time.perf_counter()
    with open(logp, "w") as log:
        p = subprocess.Popen(cmd, stdout=log, stderr=subprocess.STDOUT, start_new_session=True)
        last, stable = -1, 0
        while p.poll() is None:
            time.sleep(0.5)
            if time.perf_counter() - t0 > limit_s:
                break
            sizes = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(outdir)
                        for f in fs if f.endswith("counter_collection.csv")) \
                if os.path.isdir(outdir) else 0
            stable = stable + 1 if sizes > 0 and sizes == last else 0
            last = sizes
            if stable >= 6:  # 3 s without growth after the CSV appeared
                break
        if p.poll() is None:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
    with open(logp) as f:
        done = "pmc worker done" in f.read()
    return _pmc_rows(outdir), done, time.perf_counter() - t0, p.returncode


def pmc_measure(ncell, nband, steps=3, limit_s=120):
    """Three rocprofv3 --pmc passes over pmc_worker (FETCH_SIZE, WRITE_SIZE, SQ
    counters), each a child process under its own time limit.  Returns the
    per-dispatch means per kernel and counter, and the calibration factors
    of FETCH_SIZE / WRITE_SIZE from the device copy of the bands."""
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    base = tempfile.mkdtemp(prefix="qx_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
    stats = {}
    for name, (counters, regex) in PMC_PASSES.items():
        d = os.path.join(base, name)
        cmd = ["timeout", "-s", "KILL", str(limit_s), prof, "--pmc", *counters,
               "--kernel-include-regex", regex,
               "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.abspath(__file__), "--pmc-worker", "--ncell", str(ncell),
               "--nband", str(nband), "--steps", str(steps)]
        print(f"# counter pass {name} ({' '.join(counters)})", file=sys.stderr, flush=True)
        rows, done, sec, rc = _pmc_pass(cmd, d, limit_s)
        print(f"# counter pass {name}: {len(rows)} rows, worker finished {done}, rc {rc}, "
              f"{sec:.0f} s", file=sys.stderr, flush=True)
        if not done or not rows:
            return None, f"pass {name}: rc {rc}, worker finished {done}, {len(rows)} rows"
        for key, vals in _pmc_last(rows, steps).items():
            stats.setdefault(key, []).extend(vals)
    mean = {f"{k[0]}:{k[1]}": float(np.mean(v)) for k, v in stats.items()}
    return mean, None


def fft_counter_roof(mean, fft_avg_ms, lower_bytes, copy_bytes):
    """Traffic and VALU figures of the FFT stage from the counters: bytes =
    counter (KiB) x 1024 / calibration, the calibration being what the same
    counter reads, per byte, for the device copy of all bands (copy_bytes
    read and written)."""
    wcal = mean.get("copy:WRITE_SIZE", 0.0) * 1024 / copy_bytes
    fcal = mean.get("copy:FETCH_SIZE", 0.0) * 1024 / copy_bytes
    if wcal <= 0 or fcal <= 0:
        return None
    out = {"fetch_calibration": round(fcal, 4), "write_calibration": round(wcal, 4),
           "per_kernel_bytes": {}}
    if not (0.25 <= fcal <= 2.0 and 0.25 <= wcal <= 2.0):
        # a streaming copy of known size reads outside this range only when the
        # counters did not see the whole dispatch: no traffic figure then
        out["calibration_implausible"] = True
        return out
    tot = 0.0
    for k in ("k_zinv_f", "k_plane_f", "k_zfwd_f"):
        rd = mean.get(f"{k}:FETCH_SIZE", 0.0) * 1024 / fcal
        wr = mean.get(f"{k}:WRITE_SIZE", 0.0) * 1024 / wcal
        out["per_kernel_bytes"][k] = {"read": round(rd), "written": round(wr)}
        tot += rd + wr
    out["traffic"] = round(tot)
    out["traffic_gbs"] = round(tot / (fft_avg_ms * 1e-3) / 1e9, 1)
    out["traffic_frac"] = round(out["traffic_gbs"] / HBM_PEAK_GBS, 4)
    out["traffic_over_lower_bound"] = round(tot / lower_bytes, 3)
    waves = mean.get("k_plane_f:SQ_WAVES", 0.0)
    valu = mean.get("k_plane_f:SQ_INSTS_VALU", 0.0)
    if waves > 0:
        out["k_plane_f"] = {
            "waves": waves, "valu_instr_per_wave": round(valu / waves, 1),
            "valu_instr": valu,
            "lds_bank_conflict_over_active_lds": round(
                mean.get("k_plane_f:SQ_LDS_BANK_CONFLICT", 0.0) /
                max(mean.get("k_plane_f:SQ_ACTIVE_INST_LDS", 1.0), 1.0), 3)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncell", type=int, default=3)
    ap.add_argument("--nband", type=int, default=0)
    ap.add_argument("--no-scf", action="store_true", help="skip the SCF and MD measurements")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-kpoints", action="store_true", help="skip the k-point measurement")
    ap.add_argument("--cpu-worker", action="store_true")
    ap.add_argument("--pmc-worker", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--md-steps", type=int, default=3, help="converged MD steps of the MD run")
    args = ap.parse_args()
    natom = 8 * args.ncell ** 3
    nband = args.nband or int(2.5 * natom)  # 20 / 160 / 540 / 1280
    if args.cpu_worker:
        cpu_worker(args.ncell, args.cpu_budget)
        return
    if args.pmc_worker:
        pmc_worker(args.ncell, nband, args.steps)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs the single-GPU path only (one rank, one MI355X); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    from qxmd_amd import pw
    t_setup = time.perf_counter()
    w = Workload(args.ncell, nband)
    gs = w.gs
    ctx = pw.PlaneWave(gs, nband)
    sc = w.scf(ctx)
    ctx.put(0, w.psi0)
    assert ctx.schmidt(0, nband)
    w.psi_orth = ctx.get(0, nband)
    sc.start(None, w.occ)  # V(r) of the starting density; KB table of the cell
    ncol = ctx.info()["ncol"]
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.1f}s: Si{natom}, grid {gs.nd}, nplw {gs.nplw}, nband {nband}, "
          f"npj {ctx.npj}, ncol {ncol}", file=sys.stderr, flush=True)

    def hpsi(prec):
        return lambda: ctx.hpsi(0, 1, nband, prec)

    # warmup, then the timed region: K full H|psi> passes, mixed precision
    for _ in range(args.warmup):
        hpsi(pw.MIXED)()
        hpsi(pw.FP64)()
    # launch timers (HIP events on the library stream) run over the timed region
    ctx.timer_start(ctx.KB_GEMM)
    ctx.timer_start(ctx.FFT_STAGE)
    ms_mixed = time_region(ctx, hpsi(pw.MIXED), args.steps)
    kb_ms, kb_n = ctx.timer_stop(ctx.KB_GEMM)
    fft_ms, fft_n = ctx.timer_stop(ctx.FFT_STAGE)
    print(f"# H|psi> mixed: {ms_mixed:.3f} ms/step; KB GEMM launches {kb_n} avg "
          f"{kb_ms / max(kb_n, 1):.3f} ms, FFT-stage launches {fft_n} avg "
          f"{fft_ms / max(fft_n, 1):.3f} ms", file=sys.stderr, flush=True)
    ms_fp64 = time_region(ctx, hpsi(pw.FP64), args.steps)
    print(f"# H|psi> fp64 : {ms_fp64:.3f} ms/step", file=sys.stderr, flush=True)
    reps = max(3, args.steps // 2)
    ms_loc32 = time_region(ctx, lambda: ctx.local(0, 1, nband, pw.MIXED), reps)
    ms_loc64 = time_region(ctx, lambda: ctx.local(0, 1, nband, pw.FP64), reps)
    ms_kb = time_region(ctx, lambda: ctx.kb(0, 1, nband), reps)
    print(f"# stages: local fp32 {ms_loc32:.3f} ms, local fp64 {ms_loc64:.3f} ms, "
          f"KB fp64 {ms_kb:.3f} ms", file=sys.stderr, flush=True)

    # rooflines of the two stages of the mixed step, from the launch timers
    nplwex = gs.nplwex
    kb_launch_flops = 2.0 * ctx.npj * nplwex * nband
    kb_avg = kb_ms / max(kb_n, 1)
    fft_avg = fft_ms / max(fft_n, 1)
    fft_bands_per_launch = int(round(nband * args.steps / max(fft_n, 1)))
    lower, per_band = fft_stage_lower_bytes(gs, ncol, fft_bands_per_launch)
    kb_peak = BF16_PEAK_TFS
    kb_roof = {"bound": "mfma", "achieved": round(kb_launch_flops / (kb_avg * 1e-3) / 1e12, 3),
               "peak": kb_peak, "unit": "TFLOP/s", "traffic": None,
               "kernel": ctx.gemm_kernel_name(pw.MIXED)
               + " (KB projection / back-projection of a band batch; achieved counts the "
                 "FP32-equivalent flops, peak is the dense bf16 MFMA peak the split products run on)",
               "launch_ms": round(kb_avg, 4), "launches": kb_n,
               "algorithmic_per_launch": kb_launch_flops}
    kb_roof["frac"] = round(kb_roof["achieved"] / kb_peak, 4)
    fft_roof = {"bound": "hbm", "achieved": round(lower / (fft_avg * 1e-3) / 1e9, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                "kernel": "k_zinv_f + k_plane_f + k_zfwd_f (FP32 pencil FFT stage of a band batch)",
                "launch_ms": round(fft_avg, 4), "launches": fft_n,
                "bands_per_launch": fft_bands_per_launch,
                "algorithmic_per_launch": lower, "algorithmic_per_band": per_band,
                "model": "lower bound of this design: sphere psi in + H psi out (FP64), FP32 psi "
                         "copy out + FP32 KB term in, z-pencil T1/T2 written and read once "
                         "(FP32 complex), V(r) once per launch"}
    fft_roof["frac"] = round(fft_roof["achieved"] / HBM_PEAK_GBS, 4)
    if not args.no_pmc:
        mean, err = pmc_measure(args.ncell, nband)
        if mean:
            cr = fft_counter_roof(mean, fft_avg, lower, 8.0 * nband * gs.nplwex)
            mf = {}
            for key, tag in (("bf3_tn", "kb_bf3_projection"), ("bf3_nn", "kb_bf3_back_projection"),
                             ("f64_tn", "kb_fp64_projection"), ("f64_nn", "kb_fp64_back_projection")):
                u = mean.get(f"{key}:MfmaUtil")
                if u is None:
                    continue
                fl = mean.get(f"{key}:MfmaFlopsBF16", 0.0) + mean.get(f"{key}:MfmaFlopsF64", 0.0)
                mf[tag] = {"mfma_util_pct": round(u, 2), "mfma_flops_per_dispatch": fl,
                           "over_algorithmic": round(fl / kb_launch_flops, 3)}
            if mf:
                kb_roof["counters"] = mf
            if cr:
                fft_roof["traffic"] = cr.pop("traffic", None)
                fft_roof["counters"] = cr
                kp = cr.get("k_plane_f")
                if kp:
                    # VALU issue roofline of the plane kernel: one wave64 VALU
                    # instruction per cycle per CU (four 16-lane SIMDs, 4 cycles
                    # each); its share of the stage time from the launch timer is
                    # not separated, so the stage time bounds it from above
                    issue = 256 * 2.4e9
                    kp["valu_issue_frac_of_stage"] = round(kp["valu_instr"] / (fft_avg * 1e-3)
                                                           / issue, 4)
        else:
            fft_roof["counters_error"] = err
        print(f"# FFT-stage counters: {fft_roof.get('counters', err)}", file=sys.stderr,
              flush=True)
    roof, other = (fft_roof, kb_roof) if fft_ms >= kb_ms else (kb_roof, fft_roof)
    order = ("bound", "achieved", "peak", "unit", "frac", "traffic")
    roof = {k: roof[k] for k in order} | {k: roof[k] for k in roof if k not in order}

    extra = {}
    if not args.no_scf:
        res = {}
        for name, prec, batched in (("fp64_batched", pw.FP64, True),
                                    ("mixed_batched", pw.MIXED, True),
                                    ("fp64_band_by_band", pw.FP64, False)):
            # one untimed step first (first-use set-up of a shape), then the timed one
            for timed in (False, True):
                ctx.put(0, w.psi_orth)
                ctx.sync()
                t0 = time.perf_counter()
                eig, _, ok = ctx.scf_band_step(nband, w.occ, gs.volume, prec=prec,
                                               batched=batched, want_rho=False)
                dt = time.perf_counter() - t0
            res[name] = (dt, eig, ok, ctx.scf_phase_ms())
            print(f"# SCF band step {name}: {dt * 1e3:.1f} ms ok={ok} phases {res[name][3]}",
                  file=sys.stderr, flush=True)
        extra["scf_step_ms"] = {k: round(v[0] * 1e3, 2) for k, v in res.items()}
        extra["scf_phase_ms"] = {k: v[3] for k, v in res.items()}
        # converged SCF from identical starts (1e-6 Ha/atom on the energy change)
        conv = {}
        for name, prec, batched, refine in (
                ("fp64_batched", pw.FP64, True, None),
                ("mixed_batched_fp64_refined", pw.MIXED, True, 1e-6),
                ("mixed_batched", pw.MIXED, True, None),
                ("fp64_band_by_band", pw.FP64, False, None)):
            conv[name] = scf_converge(ctx, w, prec, batched, refine)
            print(f"# SCF to 1e-6 Ha/atom {name}: {conv[name]['ms']:.0f} ms, "
                  f"{conv[name]['iterations']} iterations "
                  f"({conv[name]['fp64_refinement_steps']} FP64 refinement)",
                  file=sys.stderr, flush=True)
        ref = conv["fp64_batched"]
        extra["scf_converged"] = {
            k: {kk: v[kk] for kk in ("ms", "iterations", "converged", "fp64_refinement_steps")}
            | {"energy_ha_per_atom": v["energy_ry"] / 2 / natom}
            for k, v in conv.items()}
        extra["scf_converged_errors_vs_fp64_batched"] = {
            k: scf_errors(conv[k], ref, natom) for k in conv if k != "fp64_batched"}
        extra["scf_converged_speedup_vs_fp64_batched"] = {
            k: round(ref["ms"] / conv[k]["ms"], 3) for k in conv if k != "fp64_batched"}
        extra["md_parts"] = md_parts(ctx, w, sc)
        print(f"# MD-step pieces: {extra['md_parts']}", file=sys.stderr, flush=True)
        extra["sbalinaspc_ms"] = extrapolation(ctx, w.psi_orth, nband)
        print(f"# sbalinaspc: {extra['sbalinaspc_ms']}", file=sys.stderr, flush=True)
        extra["md_step"] = md_run(ctx, w, nsteps=args.md_steps)
        print(f"# converged MD step: {extra['md_step']}", file=sys.stderr, flush=True)
    if not args.no_kpoints:
        extra["kpoints"] = kpoint_bench(args.steps)
        print(f"# k-points: {extra['kpoints']}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        d = cpu_baseline(args.ncell, threads)
        model = cpu_model()
        if d:
            cpu = {"value": round(d["bands"] / d["seconds"], 3), "unit": "band-applies/s",
                   "cores": threads, "kind": "port",
                   "sample": f"oracle FP64 H|psi> (C, OpenMP, textbook FFT) on {d['bands']} "
                             f"bands of the same Si{natom} grid and sphere, {d['seconds']:.1f} s",
                   "cpu_model": model, "nproc": os.cpu_count()}
        d1 = cpu_baseline(args.ncell, 1, budget_s=6.0)
        if d1:
            extra["cpu_baseline_1core"] = {
                "value": round(d1["bands"] / d1["seconds"], 3), "unit": "band-applies/s",
                "cores": 1, "kind": "port", "cpu_model": model,
                "sample": f"oracle FP64 H|psi> (C) on {d1['bands']} bands, {d1['seconds']:.1f} s"}

    value = nband / (ms_mixed * 1e-3)
    line = {
        "metric": "H|psi> band-applies/s + SCF step time, Si216 mixed vs FP64 at 1e-6 Ha/atom",
        "value": round(value, 1),
        "unit": "band-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 FFT + bf16x3 split-product KB (mixed); fp64 dense",
        "data": "synthetic (seeded Si cell and bands; the reference's Si LDA pseudopotential)",
        "config": {"workload": f"Si{natom} Gamma H|psi>, {nband} bands, grid {gs.nd[0]}^3, "
                               f"nplw {gs.nplw}, {ctx.npj} KB projectors",
                   "natom": natom, "nband": nband, "grid": list(gs.nd), "nplw": gs.nplw,
                   "npj": ctx.npj, "ecut_ry": 20.0, "ecutdens_ry": 80.0},
        "fp64_band_applies_per_s": round(nband / (ms_fp64 * 1e-3), 1),
        "fp64_ms_per_step": round(ms_fp64, 4),
        "stage_ms": {"local_fp32": round(ms_loc32, 4), "local_fp64": round(ms_loc64, 4),
                     "kb_fp64": round(ms_kb, 4),
                     "kb_gemm_mixed_launches": round(kb_ms / args.steps, 4)},
        "roofline": roof,
        "roofline_other": other,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
