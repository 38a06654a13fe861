"""Host-side mirror of the reference's plane-wave path interface.

``GSphere`` runs the set-up the reference does in reclat_ini/strvec/reclat
(grid rule, sorted G vectors, FFT maps) through the library; ``PlaneWave``
owns one device context (G sphere + FFT grid + potential + KB tables + band
storage) and exposes the reference's operations by their QXMD names.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import QxError, check, f64, i32, load, track

FP64, MIXED = 0, 1
GEMM_OWN, GEMM_LIB, GEMM_OWN_F32 = 0, 1, 2


def set_gemm_path(path):
    """Kernels of the sizable contractions, process-wide: GEMM_OWN (ours;
    split bf16x3 products in the mixed form), GEMM_LIB (hipBLASLt, a
    cross-check), GEMM_OWN_F32 (ours, plain FP32 MFMA in the mixed form)."""
    check(load().qx_set_gemm_path(int(path)), "qx_set_gemm_path")


def gemm_path():
    return int(load().qx_gemm_path())


SCHED_TIMED, SCHED_STREAMK, SCHED_SLICED = 0, 1, 2


def set_gemm_schedule(mode):
    """Work schedules of our contraction kernels, process-wide: SCHED_TIMED
    (each shape takes the faster of stream-K and K-sliced on first use),
    SCHED_STREAMK or SCHED_SLICED only."""
    check(load().qx_set_gemm_schedule(int(mode)), "qx_set_gemm_schedule")


def gemm_schedule():
    return int(load().qx_gemm_schedule())


def clmrcp(b, dkmax2):
    """The Ewald reciprocal-lattice list (clmrcp, Sources/bulk.f90:556) for
    reciprocal vectors b (3, 3; columns, without 2 pi) and |G|^2 <= dkmax2:
    dict of nga, ngb, ngc, gx, gy, gz, recnrm (m = 1..numk2, the reference's
    order) and kcsma (the largest |h|, |k|, |l|)."""
    L = load()
    bf = f64(np.asarray(b, np.float64).T.ravel())  # column-major
    kc = np.zeros(3, np.int32)
    z = np.zeros(1, np.int32)
    zd = np.zeros(1)
    n = -L.qx_clmrcp(bf, float(dkmax2), 0, z, z, z, zd, zd, zd, zd, kc)
    if n <= 0:
        raise QxError(f"qx_clmrcp failed ({-n})")
    nga, ngb, ngc = (np.zeros(n, np.int32) for _ in range(3))
    gx, gy, gz, rec = (np.zeros(n) for _ in range(4))
    m = L.qx_clmrcp(bf, float(dkmax2), n, nga, ngb, ngc, gx, gy, gz, rec, kc)
    if m != n:
        raise QxError("qx_clmrcp failed")
    return dict(nga=nga, ngb=ngb, ngc=ngc, gx=gx, gy=gy, gz=gz, recnrm=rec, kcsma=kc)


def ksphere(bvec, kfrac, ecut, nd):
    """The sphere |k+G|^2 <= ecut of k-point kfrac (units of the reciprocal
    vectors bvec, columns with 2 pi) on grid nd, in the library's storage
    order: (hkl (npw, 3) int32, |k+G|^2 (npw,))."""
    L = load()
    b = f64(np.asarray(bvec, np.float64).T.ravel())
    kf = f64(np.asarray(kfrac, np.float64))
    n = -L.qx_ksphere(b, kf, float(ecut), i32(nd), 0, None, None)
    if n <= 0:
        raise QxError("qx_ksphere failed")
    hkl = np.zeros((n, 3), np.int32)
    kin = np.zeros(n)
    if L.qx_ksphere(b, kf, float(ecut), i32(nd), n, hkl.ctypes.data_as(C.c_void_p),
                    kin.ctypes.data_as(C.c_void_p)) != n:
        raise QxError("qx_ksphere failed")
    return hkl, kin


class GSphere:
    """get_mesh + strvec + set_rciprl + reclat for cell ``hcell`` (columns are
    lattice vectors, bohr) and cutoffs in Ry."""

    def __init__(self, hcell, ecut=20.0, ecutdens=80.0):
        L = load()
        h = f64(np.asarray(hcell, np.float64).T.ravel())  # column-major
        nd = np.zeros(3, np.int32)
        check(L.qx_get_mesh(h, float(ecutdens), nd), "qx_get_mesh")
        self.nd = tuple(int(v) for v in nd)
        counts = np.zeros(6, np.int32)
        cap = 1
        rc = L.qx_strvec(h, float(ecutdens), float(ecut), cap, counts, np.zeros(1, np.int32),
                         np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1), None, None,
                         None)
        cap = -rc
        nga, ngb, ngc = (np.zeros(cap, np.int32) for _ in range(3))
        rec, gx, gy, gz = (np.zeros(cap) for _ in range(4))
        num = L.qx_strvec(h, float(ecutdens), float(ecut), cap, counts, nga, ngb, ngc, rec,
                          gx.ctypes.data_as(C.c_void_p), gy.ctypes.data_as(C.c_void_p),
                          gz.ctypes.data_as(C.c_void_p))
        if num <= 0:
            raise QxError("qx_strvec failed")
        self.num, self.nplw5, self.nplw, self.nplw2, self.nplwc, self.nplw7 = (int(x) for x in counts)
        self.nga, self.ngb, self.ngc, self.recnrm = nga, ngb, ngc, rec
        self.gx, self.gy, self.gz = gx, gy, gz
        self.nplwex = 2 * (self.nplw + 1)
        self.N = int(np.prod(self.nd))
        self.hcell = np.asarray(hcell, np.float64)
        self.volume = float(abs(np.linalg.det(self.hcell)))

    def ijkgd(self):
        out = np.zeros(2 * self.nplw5 + 1, np.int32)
        check(load().qx_ijkgd(self.nplw5, self.nga, self.ngb, self.ngc, i32(self.nd), out),
              "qx_ijkgd")
        return out

    def dkgnrm(self):
        d = np.repeat(self.recnrm[: self.nplw + 1], 2)
        return d


class PlaneWave:
    """Device context for one G sphere (slots 0 = psi, 1 = H psi)."""

    def __init__(self, gs: GSphere, nbmax: int):
        L = load()
        self.gs = gs
        self.nbmax = int(nbmax)
        h = C.c_void_p()
        n1, n2, n3 = gs.nd
        check(L.qx_create(C.byref(h), n1, n2, n3, gs.nplw, gs.nga, gs.ngb, gs.ngc, gs.recnrm,
                          self.nbmax), "qx_create")
        self.h = h
        self.npj = 0
        track(self)

    def close(self):
        if self.h:
            load().qx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- boundary data
    def set_glocal(self, glocal):
        check(load().qx_set_glocal(self.h, f64(glocal)), "qx_set_glocal")

    def set_kb(self, eqlmr, lmx_atom, vclalp_atom):
        """eqlmr: (natom, lmmax, nplwex) C-order == Fortran eqlmr_(nplwex, lmmax, natom)."""
        e = f64(eqlmr)
        natom, lmmax, n = e.shape
        assert n == self.gs.nplwex
        check(load().qx_set_kb(self.h, e, lmmax, natom, i32(lmx_atom), f64(vclalp_atom)),
              "qx_set_kb")
        self.natom_kb = natom
        info = np.zeros(5, np.int32)
        load().qx_info(self.h, info)
        self.npj = int(info[3])

    # ---- KB inputs built on the device (stscfn_pw, setnlc_g)
    def set_nlpp_tables(self, t):
        """Register the nonlocal tables (qxmd_amd.nlpp.NlppTables) setqlm reads."""
        ntype, lmmax = t.itola.shape
        self._nl_lmmax = lmmax
        qr, qi = f64(t.qlmr), f64(t.qlmi)
        check(load().qx_set_nlpp_tables(
            self.h, ntype, lmmax, t.tbqlm.shape[1], t.tbqlm.shape[2] - 1, i32(t.lmx), f64(t.clalp),
            f64(t.tbqlm), f64(t.tbqlma), f64(t.dlqlm), i32(t.itola), i32(t.itolm), i32(t.itolmf),
            i32(t.litoll), qr.shape[0], qr.shape[1] - 1, qr, qi), "qx_set_nlpp_tables")

    def stscfn_pw(self, ratm, hcell, nd1v, nd1vks, nga, ngb, ngc, nplwcs, kmax, lclust=False,
                  copy_out=True):
        """Structure factors on the device (kept for setnlc_g); ratm (natom, 3).
        Returns ycos, ysin (natom, nplwcs + 1) when copy_out."""
        r = f64(ratm)
        natom = r.shape[0]
        ga = i32(nga)
        yc = np.zeros((natom, nplwcs + 1)) if copy_out else None
        ys = np.zeros((natom, nplwcs + 1)) if copy_out else None
        check(load().qx_stscfn_pw(
            self.h, r, natom, f64(np.asarray(hcell).T), i32(nd1v), i32(nd1vks), int(bool(lclust)),
            int(kmax[0]), int(kmax[1]), int(kmax[2]), int(nplwcs), len(ga) - 1, ga, i32(ngb),
            i32(ngc), yc.ctypes.data_as(C.c_void_p) if copy_out else None,
            ys.ctypes.data_as(C.c_void_p) if copy_out else None), "qx_stscfn_pw")
        return (yc, ys) if copy_out else None

    def setnlc_g(self, ntype, nhk1, nhk2, natom, rvol, ycos=None, ysin=None, lkbppi=None,
                 lking=None, nplwcs=None):
        """KB table on the device from the registered tables and the structure
        factors (the device copy of the last stscfn_pw when ycos is None).
        Returns vclalp (ntype, lmmax)."""
        lk = np.ones(ntype, np.int32) if lkbppi is None else i32(lkbppi)
        lg = np.zeros(ntype, np.int32) if lking is None else i32(lking)
        if ycos is not None:
            yc, ys = f64(ycos), f64(ysin)
            nplwcs = yc.shape[1] - 1
            pc, ps = yc.ctypes.data_as(C.c_void_p), ys.ctypes.data_as(C.c_void_p)
        else:
            pc = ps = None
            nplwcs = self.gs.nplw if nplwcs is None else nplwcs
        vcl = np.zeros((ntype, getattr(self, "_nl_lmmax", 0) or 1))
        check(load().qx_setnlc_g(self.h, self.gs.nplw, ntype, i32(nhk1), i32(nhk2), natom, lk, lg,
                                 pc, ps, int(nplwcs), float(rvol),
                                 vcl.ctypes.data_as(C.c_void_p)), "qx_setnlc_g")
        self.natom_kb = natom
        info = np.zeros(5, np.int32)
        load().qx_info(self.h, info)
        self.npj = int(info[3])
        return vcl

    def kb_forces(self, slot, nb, occ, fnlc=None):
        """KB Hellmann-Feynman forces of the bands in slot (kbnlpp_frc2): added to
        fnlc (natom, 3), zeros if None; returns it."""
        natom = self.natom_kb
        f = np.zeros((natom, 3)) if fnlc is None else f64(fnlc).copy()
        nw = self.gs.nplw + 1
        check(load().qx_kbnlpp_frc2_slot(self.h, slot, nb, f64(occ), f64(self.gs.gx[:nw]),
                                         f64(self.gs.gy[:nw]), f64(self.gs.gz[:nw]), f),
              "qx_kbnlpp_frc2_slot")
        return f

    def get_kb(self, natom, lmmax):
        """KB table in the reference layout: eqlmr (natom, lmmax, nplwex), vclalp (natom, lmmax)."""
        eq = np.zeros((natom, lmmax, self.gs.nplwex))
        vcl = np.zeros((natom, lmmax))
        check(load().qx_get_kb(self.h, eq, vcl.ctypes.data_as(C.c_void_p)), "qx_get_kb")
        return eq, vcl

    @staticmethod
    def gemm_kernel_name(prec):
        """What runs the large contractions (KB, overlap, rotation) at precision prec."""
        path = load().qx_gemm_path()
        if path == GEMM_LIB:
            return ("hipBLASLt FP32 GEMM, FAST_TF32 split products on bf16 MFMA"
                    if prec == MIXED else "hipBLASLt FP64 GEMM")
        if prec == MIXED:
            return ("k_mg_ps3: split bf16x3 products on v_mfma_f32_16x16x32_bf16, operands "
                    "split beforehand (the KB table once, psi by the z-inverse pass)"
                    if path == GEMM_OWN else "k_gemm_big<float>: v_mfma_f32_16x16x4f32")
        return "k_mg_f64: v_mfma_f64_16x16x4f64, 128 or 96 tile edges per shape"

    def set_cg_graph(self, on):
        """Replay the batched CG's captured iteration graph (on by default)."""
        check(load().qx_set_cg_graph(self.h, int(bool(on))), "qx_set_cg_graph")

    def cg_graph_count(self):
        return int(load().qx_cg_graph_count(self.h))

    def info(self):
        info = np.zeros(5, np.int32)
        load().qx_info(self.h, info)
        return dict(ncol=int(info[0]), npair=int(info[1]), nH=int(info[2]), npj=int(info[3]),
                    nplwex=int(info[4]))

    # ---- band storage
    def put(self, slot, psi):
        p = f64(psi)
        check(load().qx_slot_put(self.h, int(slot), p, p.shape[0]), "qx_slot_put")

    def get(self, slot, nb):
        out = np.zeros((nb, self.gs.nplwex))
        check(load().qx_slot_get(self.h, int(slot), out, nb), "qx_slot_get")
        return out

    def sync(self):
        check(load().qx_sync(self.h), "qx_sync")

    # ---- Bloch bands at several k-points (k-slots, complex interleaved)
    def kpoints_set(self, kfrac, bvec, ecut, nb):
        """k-points kfrac (nk, 3) in units of the reciprocal vectors bvec
        (columns, 2 pi included), sphere |k+G|^2 <= ecut (Ry), nb bands each.
        Returns the sphere sizes (nk,)."""
        kf = f64(np.asarray(kfrac, np.float64).reshape(-1, 3))
        nk = kf.shape[0]
        npw = np.zeros(nk, np.int32)
        check(load().qx_kpoints_set(self.h, nk, kf, f64(np.asarray(bvec).T), float(ecut), int(nb),
                                    npw), "qx_kpoints_set")
        self._knpw = npw.copy()
        self._knb = int(nb)
        return npw

    def kpoints_sphere(self, ik):
        """(hkl (npw, 3), |k+G|^2 (npw,)) of k-point ik, in storage order."""
        n = int(self._knpw[ik])
        hkl = np.zeros((n, 3), np.int32)
        kin = np.zeros(n)
        check(load().qx_kpoints_sphere(self.h, int(ik), hkl, kin), "qx_kpoints_sphere")
        return hkl, kin

    def kput(self, slot, ik, psi):
        """psi (nb, npw) complex -> k-slot."""
        p = np.asarray(psi, np.complex128)
        x = f64(p.view(np.float64).reshape(p.shape[0], -1))
        check(load().qx_kslot_put(self.h, int(slot), int(ik), x, p.shape[0]), "qx_kslot_put")

    def kget(self, slot, ik, nb=None):
        nb = self._knb if nb is None else int(nb)
        out = np.zeros((nb, 2 * int(self._knpw[ik])))
        check(load().qx_kslot_get(self.h, int(slot), int(ik), out, nb), "qx_kslot_get")
        return out.view(np.complex128)

    def kset_kb(self, ik, beta, D):
        """beta (npj, 2 npw_ik) complex interleaved, D (npj,)."""
        b = f64(beta)
        check(load().qx_kset_kb(self.h, int(ik), b.shape[0], b, f64(D)), "qx_kset_kb")

    def khpsi(self, src=0, dst=1, nb=None, per_band=False):
        nb = self._knb if nb is None else int(nb)
        check(load().qx_khpsi_slot(self.h, int(src), int(dst), nb, int(bool(per_band))),
              "qx_khpsi_slot")

    def ksolve(self, src=0, dst=1, max_iter=100, tol=1e-10):
        """Lowest bands of every k-point in the fixed potential (block
        Rayleigh-Ritz from the bands of k-slot src).  Returns (eig (nk, nb),
        steps)."""
        nk = len(self._knpw)
        eig = np.zeros((nk, self._knb))
        it = np.zeros(1, np.int32)
        check(load().qx_ksolve_slot(self.h, int(src), int(dst), int(max_iter), float(tol), eig, it),
              "qx_ksolve_slot")
        return eig, int(it[0])

    # ---- operations on resident slots
    def slot_copy(self, src, dst, nb):
        check(load().qx_slot_copy(self.h, int(src), int(dst), int(nb)), "qx_slot_copy")

    def hpsi(self, src=0, dst=1, nb=None, prec=FP64):
        check(load().qx_hpsi_slot(self.h, src, dst, nb, prec), "qx_hpsi_slot")

    def local(self, src, dst, nb, prec=FP64, accumulate=False):
        check(load().qx_local_slot(self.h, src, dst, nb, prec, int(accumulate)), "qx_local_slot")

    def kb(self, src, dst, nb):
        check(load().qx_kb_slot(self.h, src, dst, nb, None), "qx_kb_slot")

    def schmidt(self, slot, nb):
        ok = C.c_int(0)
        check(load().qx_schmidt_slot(self.h, slot, nb, C.byref(ok)), "qx_schmidt_slot")
        return bool(ok.value)

    def rotate(self, psi, hpsi, nb):
        eig = np.zeros(nb)
        check(load().qx_rotate_slot(self.h, psi, hpsi, nb, eig), "qx_rotate_slot")
        return eig

    def density(self, slot, nb, occ, rvol, prec=FP64):
        rho = np.zeros(self.gs.N)
        check(load().qx_density_slot(self.h, slot, nb, f64(occ), float(rvol),
                                     rho.ctypes.data_as(C.c_void_p), prec), "qx_density_slot")
        return rho

    def ekib(self, slot, nb):
        out = np.zeros(nb)
        check(load().qx_ekib_slot(self.h, slot, nb, out), "qx_ekib_slot")
        return out

    def eigen_bkl(self, nb, ekib, eig, noband=0, itermx=4, iteremx=4, tolera=1e-10, threinn=0.0,
                  wdinner=0.9, roughfzan=0.3, roughfeig=0.3, prec=FP64, batched=False):
        e = f64(eig).copy()
        bz = np.zeros(nb)
        nr = np.zeros(nb, np.int32)
        L = load()
        if batched:
            check(L.qx_eigen_bkl_batched_slot(self.h, 0, 1, nb, f64(ekib), e, noband, itermx,
                                              iteremx, tolera, roughfzan, roughfeig, bz, nr, prec),
                  "qx_eigen_bkl_batched_slot")
        else:
            check(L.qx_eigen_bkl_slot(self.h, 0, 1, nb, f64(ekib), e, noband, itermx, iteremx,
                                      tolera, threinn, wdinner, roughfzan, roughfeig, bz, nr, prec),
                  "qx_eigen_bkl_slot")
        return e, bz, nr

    def scf_band_step(self, nb, occ, rvol, noband=0, prec=FP64, batched=False, want_rho=True):
        eig = np.zeros(nb)
        rho = np.zeros(self.gs.N) if want_rho else None
        ok = C.c_int(0)
        check(load().qx_scf_band_step_ex(self.h, 0, 1, nb, f64(occ), noband, float(rvol), eig,
                                         rho.ctypes.data_as(C.c_void_p) if rho is not None else None,
                                         C.byref(ok), prec, int(batched)), "qx_scf_band_step")
        return eig, rho, bool(ok.value)

    # ---- timing
    def scf_phase_ms(self):
        """Phase times (ms) of the last SCF band step."""
        ms = (C.c_float * 5)()
        check(load().qx_scf_phase_ms(self.h, ms), "qx_scf_phase_ms")
        return dict(zip(("energy", "rotation", "cg", "schmidt", "density"),
                        (round(float(x), 3) for x in ms)))

    KB_GEMM, FFT_STAGE = 0, 1

    def timer_start(self, tid):
        check(load().qx_launch_timer(self.h, tid, 1, None, None), "qx_launch_timer")

    def timer_stop(self, tid):
        """(summed ms, launches) of launch timer tid since timer_start."""
        ms, n = C.c_double(0), C.c_int(0)
        check(load().qx_launch_timer(self.h, tid, 0, C.byref(ms), C.byref(n)), "qx_launch_timer")
        return float(ms.value), int(n.value)

    def record(self, slot):
        check(load().qx_event_record(self.h, slot), "qx_event_record")

    def elapsed_ms(self, s0, s1):
        ms = C.c_float(0)
        check(load().qx_event_elapsed(self.h, s0, s1, C.byref(ms)), "qx_event_elapsed")
        return float(ms.value)

    # ---- Fortran-shaped drop-ins on host arrays
    def hcnonloc(self, cgjr, rhcr):
        r = f64(rhcr).copy()
        check(load().qx_hcnonloc(self.h, f64(cgjr), r), "qx_hcnonloc")
        return r

    def hckbpp_g(self, cgjr, rhcr, want_slmir=False, lmmax=None, natom=None):
        c = f64(cgjr)
        r = f64(rhcr).copy()
        nb = c.shape[0]
        sl = np.zeros((nb, natom, lmmax)) if want_slmir else None
        check(load().qx_hckbpp_g(self.h, c, r, nb,
                                 sl.ctypes.data_as(C.c_void_p) if sl is not None else None),
              "qx_hckbpp_g")
        return (r, sl) if want_slmir else r

    def schmidt_host(self, cgjr):
        c = f64(cgjr).copy()
        ok = C.c_int(0)
        check(load().qx_schmidt(self.h, c, c.shape[0], C.byref(ok)), "qx_schmidt")
        return c, bool(ok.value)

    def rseigqr(self, a):
        """Eigen-decomposition of a symmetric matrix on the device (FP64)."""
        n = a.shape[0]
        m = np.asfortranarray(np.asarray(a, np.float64)).copy(order="F")
        w = np.zeros(n)
        check(load().qx_rseigqr(self.h, m.ctypes.data_as(C.c_void_p), n, w), "qx_rseigqr")
        return w, np.asarray(m)

    def icholesky(self, s):
        n = s.shape[0]
        m = np.asfortranarray(np.asarray(s, np.float64)).copy(order="F")
        ok = C.c_int(0)
        check(load().qx_icholesky(self.h, m.ctypes.data_as(C.c_void_p), n, C.byref(ok)),
              "qx_icholesky")
        return np.asarray(m), bool(ok.value)

    # ---- local-pseudopotential forces (flocal_g)
    def set_lopp(self, svco):
        """svco (ntype, nplw5 + 1): local form factors over the density sphere."""
        gs = self.gs
        sv = f64(svco)
        n5 = gs.nplw5
        assert sv.shape[1] == n5 + 1
        check(load().qx_set_lopp(self.h, sv.shape[0], n5, sv, f64(gs.gx[: n5 + 1]),
                                 f64(gs.gy[: n5 + 1]), f64(gs.gz[: n5 + 1])), "qx_set_lopp")

    def flocal_g(self, ntype, nhk1, nhk2, iatmpt, llking, natom, rhgr, kmax, tables, floc=None):
        """Host-array flocal_g: tables = [(DCQ_a, DSQ_a)] with (natom, 2 kmax_a + 1)."""
        gs = self.gs
        n5 = gs.nplw5
        f = np.zeros((natom, 3)) if floc is None else f64(floc).copy()
        t = [(f64(c), f64(s)) for c, s in tables]
        check(load().qx_flocal_g(self.h, f, int(ntype), i32(nhk1), i32(nhk2), i32(iatmpt),
                                 int(natom), i32(llking), n5, i32(gs.nga[: n5 + 1]),
                                 i32(gs.ngb[: n5 + 1]), i32(gs.ngc[: n5 + 1]), f64(rhgr),
                                 int(kmax[0]), int(kmax[1]), int(kmax[2]), t[0][0], t[1][0],
                                 t[2][0], t[0][1], t[1][1], t[2][1]), "qx_flocal_g")
        return f

    def flocal_g_dev(self, ntype, nhk1, nhk2, iatmpt, llking, natom, floc=None):
        """Device-resident flocal_g on the context's density sphere."""
        f = np.zeros((natom, 3)) if floc is None else f64(floc).copy()
        check(load().qx_flocal_g_dev(self.h, int(ntype), i32(nhk1), i32(nhk2), i32(iatmpt),
                                     i32(llking), int(natom), f), "qx_flocal_g_dev")
        return f

    # ---- restart files of a slot (QM_eig)
    def svwfns(self, fname, slot, nb, eig, occ, lsreal8=True):
        check(load().qx_svwfns_slot(self.h, str(fname).encode(), slot, nb, f64(eig), f64(occ),
                                    int(bool(lsreal8))), "qx_svwfns_slot")

    def rdwfns(self, fname, slot, nb):
        """Returns (eig, occ, nband in the file)."""
        eig, occ = np.zeros(nb), np.zeros(nb)
        nf = C.c_int(0)
        check(load().qx_rdwfns_slot(self.h, str(fname).encode(), slot, nb, eig, occ,
                                    C.byref(nf)), "qx_rdwfns_slot")
        return eig, occ, nf.value

    # ---- wave-function extrapolation (sbalinaspc)
    def vlocal_g(self, ntype, nhk1, nhk2, iatmpt, llking, natom, rvol, accumulate=False,
                 copy_out=True):
        """Local pseudopotential into the context's external potential
        (vlocal_g on the device); returns (vext (FFT order), thrhgr) when
        copy_out, else None."""
        vx = np.zeros(self.gs.N) if copy_out else None
        th = np.zeros(2 * (self.gs.nplw5 + 1)) if copy_out else None
        check(load().qx_vlocal_g(self.h, int(ntype), i32(nhk1), i32(nhk2), i32(iatmpt),
                                 i32(llking), int(natom), float(rvol), int(bool(accumulate)),
                                 vx.ctypes.data_as(C.c_void_p) if copy_out else None,
                                 th.ctypes.data_as(C.c_void_p) if copy_out else None),
              "qx_vlocal_g")
        return (vx, th) if copy_out else None

    def aspc_init(self, nprvmx, ihest):
        check(load().qx_aspc_init(self.h, int(nprvmx), int(ihest)), "qx_aspc_init")

    def sbalinaspc(self, slot, nb, nstep, nstep_ini=0):
        """Returns (iaspc, lcholesky)."""
        info = np.zeros(2, np.int32)
        check(load().qx_sbalinaspc_slot(self.h, slot, nb, int(nstep), int(nstep_ini),
                                        info.ctypes.data_as(C.c_void_p)), "qx_sbalinaspc_slot")
        return int(info[0]), bool(info[1])

    def roteig(self, slot, nb, kpwvx):
        check(load().qx_roteig_slot(self.h, slot, nb, int(kpwvx)), "qx_roteig_slot")

    def aspc_get(self, k, nb):
        out = np.zeros((nb, self.gs.nplwex))
        check(load().qx_aspc_get(self.h, int(k), nb, out), "qx_aspc_get")
        return out

    def aspc_put(self, k, psi, keypwv):
        p = f64(psi)
        check(load().qx_aspc_put(self.h, int(k), p.shape[0], p, int(keypwv)), "qx_aspc_put")

    # ---- Ewald reciprocal-space forces (frcrec)
    def set_ewald(self, ew):
        """ew: nga, ngb, ngc, gx, gy, gz, recnrm (m = 1..numk2), recnrmex
        (0..numk2), gamma, volume."""
        n = len(ew["nga"])
        check(load().qx_set_ewald(self.h, n, i32(ew["nga"]), i32(ew["ngb"]), i32(ew["ngc"]),
                                  f64(ew["gx"]), f64(ew["gy"]), f64(ew["gz"]), f64(ew["recnrm"]),
                                  f64(ew["recnrmex"]), float(ew["gamma"]), float(ew["volume"])),
              "qx_set_ewald")

    def frcrec(self, ntype, nhk1, nhk2, zv, nel, natom, iatoit, iatmpt, tables=None, fclm=None,
               ecolmb=0.0, atom_ecolmb=None):
        """Returns (fclm (natom, 3), ecolmb, atom_ecolmb); tables = per-ion
        [(DCQ_a, DSQ_a)] (nion, 2 kcsma_a + 1), or None for the device tables."""
        f = np.zeros((natom, 3)) if fclm is None else f64(fclm).copy()
        ae = np.zeros(natom) if atom_ecolmb is None else f64(atom_ecolmb).copy()
        e = C.c_double(ecolmb)
        nion = len(iatmpt)
        L = load()
        if tables is None:
            check(L.qx_frcrec_dev(self.h, f, int(ntype), i32(nhk1), i32(nhk2), f64(zv), int(nel),
                                  C.byref(e), ae, i32(iatoit), i32(iatmpt), int(natom), nion),
                  "qx_frcrec_dev")
        else:
            t = [(f64(c), f64(s)) for c, s in tables]
            km = [(c.shape[1] - 1) // 2 for c, _ in t]
            check(L.qx_frcrec(self.h, f, int(ntype), i32(nhk1), i32(nhk2), f64(zv), int(nel),
                              C.byref(e), ae, i32(iatoit), i32(iatmpt), int(natom), nion, *km,
                              t[0][0], t[1][0], t[2][0], t[0][1], t[1][1], t[2][1]), "qx_frcrec")
        return f, e.value, ae

    # ---- SCF glue (harting, xclda, potential assembly, charge mixing)
    def scf_set_grid(self, recnrmex, vext=None, rhocore=None, rdelv=None):
        """Density sphere of the context's G list, Hartree kernel recnrmex
        (0..nplw5), external potential and core charge on the grid (FFT order)."""
        gs = self.gs
        self._ijk = gs.ijkgd()
        rd = gs.volume / gs.N if rdelv is None else float(rdelv)
        vx = None if vext is None else f64(vext)
        rc_ = None if rhocore is None else f64(rhocore)
        check(load().qx_scf_set_grid(
            self.h, gs.nplw5, self._ijk, f64(recnrmex),
            vx.ctypes.data_as(C.c_void_p) if vx is not None else None,
            rc_.ctypes.data_as(C.c_void_p) if rc_ is not None else None, rd), "qx_scf_set_grid")

    def mxchgg_alloc(self, imxchg=1, itratn=10, cmetric=20.0, wkerker=1.0, nplwc=None):
        """Charge mixer over the nplwc sphere (default: the wave-function sphere)."""
        gs = self.gs
        nplwc = gs.nplwc if nplwc is None else int(nplwc)
        dkg = f64(np.repeat(gs.recnrm[: nplwc + 1], 2))
        check(load().qx_mxchgg_alloc(self.h, 2 * (nplwc + 1), int(imxchg), int(itratn),
                                     float(cmetric), float(wkerker), float(gs.recnrm[1]),
                                     float(gs.recnrm[nplwc]), dkg, gs.volume),
              "qx_mxchgg_alloc")

    def pwpchg_slot(self, slot, nb, occ, prec=FP64):
        check(load().qx_pwpchg_slot(self.h, slot, nb, f64(occ), self.gs.volume, prec),
              "qx_pwpchg_slot")

    def outchg(self):
        """rhgr of the density the context holds (after a band step)."""
        check(load().qx_outchg(self.h), "qx_outchg")

    def pwpmxchgg(self, aslh, bslh, lreset=False, lmixhold=False):
        """Returns (kntchg, history resets)."""
        info = np.zeros(2, np.int32)
        check(load().qx_pwpmxchgg(self.h, int(bool(lreset)), float(aslh), float(bslh),
                                  int(bool(lmixhold)), info.ctypes.data_as(C.c_void_p)),
              "qx_pwpmxchgg")
        return int(info[0]), int(info[1])

    def scf_potential(self):
        """V = vhar + vext + vexc into the context; returns {ehar, elcl, eexc,
        evexc, ecor, evcor}."""
        e = np.zeros(6)
        check(load().qx_scf_potential(self.h, e), "qx_scf_potential")
        return dict(zip(("ehar", "elcl", "eexc", "evexc", "ecor", "evcor"), e))

    _SCF = {"rho": 0, "rhgr": 1, "rinr": 2, "vhar": 3, "vexc": 4, "eeexc": 5, "glocal": 6,
            "cmk": 7}

    def scf_get(self, what, n=None):
        k = self._SCF[what]
        size = 2 * (self.gs.nplw5 + 1) if k in (1, 2) else (n if k == 7 else self.gs.N)
        out = np.zeros(size)
        check(load().qx_scf_get(self.h, k, out), "qx_scf_get")
        return out

    def scf_put(self, what, a):
        check(load().qx_scf_put(self.h, self._SCF[what], f64(a)), "qx_scf_put")

    def mxchgg(self, rhgr, rinr, aslh, bslh, lreset=False, lmixhold=False):
        """Host-array mxchgg: returns (rhgr mixed, rinr, glocal = chgg2r(rhgr))."""
        gs = self.gs
        ijk = gs.ijkgd()
        rh, ri = f64(rhgr).copy(), f64(rinr).copy()
        g = np.zeros(gs.N)
        lr = C.c_int(int(bool(lreset)))
        check(load().qx_mxchgg(self.h, C.byref(lr), rh, ri, gs.nplw5, float(aslh), float(bslh), g,
                               None, gs.N, ijk, int(bool(lmixhold))), "qx_mxchgg")
        return rh, ri, g

    def softchg(self, cgjr, occ, rvol):
        c = f64(cgjr)
        rho = np.zeros(self.gs.N)
        check(load().qx_softchg(self.h, c, c.shape[0], f64(occ), float(rvol), rho), "qx_softchg")
        return rho


def rfft3(key, fft3x, fft3y, nd):
    x, y = f64(fft3x).copy(), f64(fft3y).copy()
    ierr = C.c_int(0)
    check(load().qx_rfft3(int(key), x, y, *[int(v) for v in nd], C.byref(ierr)), "qx_rfft3")
    return x, y


def fft3(key, fft3x, fft3y, nd):
    x, y = f64(fft3x).copy(), f64(fft3y).copy()
    ierr = C.c_int(0)
    check(load().qx_fft3(int(key), x, y, *[int(v) for v in nd], C.byref(ierr)), "qx_fft3")
    return x, y


def xclda(rho, rhocore=None, rdelv=1.0):
    """xclda on the device: (vexc, eeexc, [eexc, evexc, ecor, evcor])."""
    r = f64(rho)
    core = None if rhocore is None else f64(rhocore)
    v, e, e4 = np.zeros(r.size), np.zeros(r.size), np.zeros(4)
    check(load().qx_xclda(r, core.ctypes.data_as(C.c_void_p) if core is not None else None,
                          r.size, float(rdelv), v, e, e4), "qx_xclda")
    return v, e, e4


def _sphere_call(fn, grid, rhgr, nplw5, ijk, nd, *extra):
    return fn(grid, rhgr, int(nplw5), *extra, None, int(np.prod(nd)), i32(nd), i32(ijk))


def harting(rhgr, recnrmex, ijk, nd):
    """Hartree potential on the grid (FFT order); ijk = ijkgd(-nplw5:nplw5)."""
    rh = f64(rhgr)
    v = np.zeros(int(np.prod(nd)))
    n5 = rh.size // 2 - 1
    check(_sphere_call(load().qx_harting, v, rh, n5, ijk, nd, f64(recnrmex)), "qx_harting")
    return v


def chgr2g(rho, nplw5, ijk, nd):
    rh = np.zeros(2 * (int(nplw5) + 1))
    check(_sphere_call(load().qx_chgr2g, f64(rho), rh, nplw5, ijk, nd), "qx_chgr2g")
    return rh


def chgg2r(rhgr, ijk, nd):
    rh = f64(rhgr)
    rho = np.zeros(int(np.prod(nd)))
    check(_sphere_call(load().qx_chgg2r, rho, rh, rh.size // 2 - 1, ijk, nd), "qx_chgg2r")
    return rho
