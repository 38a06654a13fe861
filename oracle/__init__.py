"""CPU FP64 oracle for the QXMD Gamma-point H|psi> / band-update path.

TEST INFRASTRUCTURE ONLY.  Used by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker; the product package
``qxmd_amd`` never imports it.

The arithmetic lives in ``qxoracle.c`` (plain C, FP64, each routine citing the
reference file:line it restates); this module loads ``liboracle.so`` through
ctypes and offers numpy-shaped wrappers.  Pinning status of every function is
listed in ``PINNING`` below.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# How each oracle function is pinned (checked by tests/test_oracle_*.py).
PINNING = {
    "get_mesh": "reference get_mesh/chkftm (oracle/_ref) + grid sizes 30/60/90/120",
    "strvec": "bit-exact vs reference strvec/set_rciprl (oracle/_ref); lattice counts",
    "ijkgd": "bit-exact vs reference reclat (oracle/_ref); point-reflection invariant",
    "rnd00": "reference rnd00 (oracle/_ref); closed-form Lehmer values",
    "rfft3/fft3": "numpy.fft (fftn / irfftn) with FFTW sign and 1/N convention",
    "hcnonloc": "constant-V closed form, single plane wave, hermiticity, numpy FFT",
    "ekinetic": "reference ekinetic2 (oracle/_ref)",
    "hckbpp_g": "reference hckbpp_g / hckbpp_ib_g (oracle/_ref); GEMM identity",
    "enkbpp_g": "reference enkbpp_g2 (oracle/_ref)",
    "schmidt": "reference schmidt2 (oracle/_ref); Psi^T Psi = I",
    "hmatrix": "reference hmatrix2 (oracle/_ref)",
    "eigh": "reference RsEIGQR (oracle/_ref) and numpy.linalg.eigh",
    "untrycore": "reference untrycore (oracle/_ref)",
    "softchg": "integral = sum occ; numpy FFT of psi(r)^2",
    "prpcg": "reference prpcg (oracle/_ref)",
    "clmrcp": "bit-exact vs reference clmrcp (oracle/_ref); lattice closed form",
    "kpoint.hk_apply/hk_dense": "k = 0 on a real state equals the C Gamma H (pinned to _ref); "
                                "hermiticity; k -> k + b1 relabelling; apply = dense matrix",
    "eigen_bkl": "per-iteration trace against numpy: the 2x2 Rayleigh-Ritz eigenvalue on "
                 "span{x, h} of a dense H, the direction h = P g + beta h_prev up to the "
                 "projected span, beta = zansap/betk, h orthogonal to the updated bands, exits at "
                 "their thresholds; converged eigenvalues = numpy eigvalsh",
}


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    src = os.path.join(_HERE, "qxoracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = C.CDLL(so)
        _declare(_LIB)
    return _LIB


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i = C.c_int
_d = C.c_double
_vp = C.c_void_p


class Ham(C.Structure):
    _fields_ = [
        ("nplw", _i), ("nplwex", _i), ("n1", _i), ("n2", _i), ("n3", _i),
        ("ijkgd_c", _vp), ("glocal", _vp), ("dkgnrm", _vp), ("eqlmr", _vp),
        ("lmmax", _i), ("natom", _i), ("lmx_atom", _vp), ("vclalp_atom", _vp),
    ]


def _declare(L):
    L.qxo_rciprl.argtypes = [_dp, _dp, C.POINTER(_d)]
    L.qxo_fft_size.argtypes = [_i, _i]
    L.qxo_fft_size.restype = _i
    L.qxo_get_mesh.argtypes = [_dp, _d, _ip]
    L.qxo_strvec_capacity.argtypes = [_dp, _d]
    L.qxo_strvec_capacity.restype = _i
    L.qxo_strvec.argtypes = [_dp, _d, _d, _d, _d, _d, _i, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp,
                             _dp, C.POINTER(_d)]
    L.qxo_strvec.restype = _i
    L.qxo_ijkgd.argtypes = [_i, _ip, _ip, _ip, _ip, _ip]
    L.qxo_dkgnrm.argtypes = [_i, _dp, _dp]
    L.qxo_rnd00.argtypes = [C.POINTER(_d)]
    L.qxo_rnd00.restype = _d
    L.qxo_cfft3.argtypes = [_i, _dp, _i, _i, _i]
    L.qxo_rfft3.argtypes = [_i, _dp, _dp, _i, _i, _i]
    L.qxo_fft3.argtypes = [_i, _dp, _dp, _i, _i, _i]
    L.qxo_wvg2r.argtypes = [_dp, _i, _vp, _i, _i, _i, _dp, _dp]
    L.qxo_hcnonloc.argtypes = [_dp, _dp, _dp, _i, _vp, _i, _i, _i, _dp, _dp]
    L.qxo_ekinetic.argtypes = [_dp, _dp, _dp, _i, _i, _vp, _vp, _vp]
    L.qxo_hckbpp_g.argtypes = [_dp, _i, _i, _i, _ip, _dp, _dp, _dp, _i, _vp]
    L.qxo_xclda.argtypes = [_dp, _vp, _i, _d, _dp, _dp, _dp]
    L.qxo_harting.argtypes = [_dp, _i, _dp, _vp, _i, _i, _i, _dp]
    L.qxo_chgr2g.argtypes = [_dp, _i, _vp, _i, _i, _i, _dp]
    L.qxo_chgg2r.argtypes = [_dp, _i, _vp, _i, _i, _i, _dp]
    L.qxo_invers.argtypes = [_dp, _dp, _i, _i]
    L.qxo_invers.restype = _i
    L.qxo_mxchgg_pulay.argtypes = [_dp, _dp, _i, _i, _dp, _d, _d, _d, _d, _d, _i, _i, _i,
                                   C.POINTER(_i), _dp, _dp, _dp, _dp]
    L.qxo_mxchgg_pulay.restype = _i
    L.qxo_mxchgg_simple.argtypes = [_dp, _dp, _i, _d, _i]
    L.qxo_roteig.argtypes = [_dp, _dp, _i, _i, _i, _i, _vp]
    L.qxo_roteig.restype = _i
    L.qxo_coef_aspc.argtypes = [_i, _dp]
    L.qxo_sbalinaspc.argtypes = [_dp, _i, _i, _dp, _i, _dp, _i, _i, _i, _i, C.POINTER(_i)]
    L.qxo_sbalinaspc.restype = _i
    L.qxo_frcrec.argtypes = [_i, _ip, _ip, _dp, _i, _i, _i, _ip, _ip, _i, _ip, _ip, _ip, _dp, _dp,
                             _dp, _dp, _dp, _d, _d, _ip, C.POINTER(_vp), C.POINTER(_vp), _dp,
                             C.POINTER(_d), _dp]
    L.qxo_flocal_g.argtypes = [_i, _ip, _ip, _ip, _ip, _i, _i, _ip, _ip, _ip, _dp, _dp, _dp, _dp,
                               _dp, _ip, C.POINTER(_vp), C.POINTER(_vp), _dp]
    L.qxo_vlocal_g.argtypes = [_i, _ip, _ip, _ip, _ip, _i, _i, _ip, _ip, _ip, _dp, _d, _ip,
                               C.POINTER(_vp), C.POINTER(_vp), _dp]
    L.qxo_vlocal_g.restype = _d
    L.qxo_kbnlpp_frc2.argtypes = [_dp, _i, _i, _i, _ip, _dp, _dp, _i, _dp, _dp, _dp, _dp, _dp,
                                  _dp]
    L.qxo_stscfn_pw.argtypes = [_dp, _dp, _ip, _ip, _i, _i, _i, _i, _i, _i, _i, _ip, _ip, _ip,
                                _dp, _dp]
    L.qxo_setnlc_g.argtypes = [_i, _i, _ip, _ip, _i, _ip, _ip, _dp, _dp, _i, _d, _i, _i, _i, _ip,
                               _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _i, _dp, _dp, _dp, _dp,
                               _dp]
    L.qxo_enkbpp_g.argtypes = [_dp, _i, _i, _ip, _dp, _dp, _i, _d]
    L.qxo_enkbpp_g.restype = _d
    L.qxo_rsdecmp.argtypes = [_dp, _i, _i]
    L.qxo_rsinvcd.argtypes = [_dp, _i, _i, _dp]
    L.qxo_schmidt.argtypes = [_dp, _i, _i, _dp]
    L.qxo_schmidt.restype = _i
    L.qxo_hmatrix.argtypes = [_dp, _dp, _i, _i, _dp, _dp]
    L.qxo_eigh.argtypes = [_dp, _i, _i, _dp]
    L.qxo_untrycore.argtypes = [_dp, _i, _i, _dp, _i, _dp]
    L.qxo_softchg.argtypes = [_dp, _i, _i, _i, _dp, _d, _vp, _i, _i, _i, _dp]
    L.qxo_prpcg.argtypes = [_d, _dp, _i, _dp]
    L.qxo_hpsi.argtypes = [C.POINTER(Ham), _dp, _dp, _i, _vp]
    L.qxo_eigen_bkl.argtypes = [C.POINTER(Ham), _dp, _dp, _dp, _dp, _i, _i, _i, _i, _d, _d, _d,
                                _d, _d, _dp, _ip]
    L.qxo_bkl_trace.argtypes = [_vp, _vp, _i]
    L.qxo_bkl_trace.restype = None
    L.qxo_bkl_trace_count.argtypes = []
    L.qxo_bkl_trace_count.restype = _i
    L.qxo_scf_band_step.argtypes = [C.POINTER(Ham), _dp, _dp, _dp, _dp, _i, _i, _d, _vp,
                                    C.POINTER(_d), C.POINTER(_d), _vp]
    L.qxo_scf_band_step.restype = _i


def _f(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- grid / sphere

def rciprl(a):
    """a: (3,3) with lattice vectors as columns, i.e. a[:, i] = vector i."""
    af = _f(np.asarray(a, np.float64).T.ravel())  # column-major storage
    b = np.zeros(9)
    v = _d()
    lib().qxo_rciprl(af, b, C.byref(v))
    return b.reshape(3, 3).T, v.value


def fft_size(n, first_dim=True):
    return lib().qxo_fft_size(int(n), 1 if first_dim else 0)


def get_mesh(hcell, ecutdens):
    nd = np.zeros(3, np.int32)
    lib().qxo_get_mesh(_f(np.asarray(hcell, np.float64).T.ravel()), float(ecutdens), nd)
    return tuple(int(x) for x in nd)


def strvec(rba, ecutdens, ecut, ecutsoft=None, ecutc=None, ecutorth=None):
    """G-vector generation and stable |G| sort.  Returns a dict with the counts
    and the sorted arrays (index 0 = G=0)."""
    ecutsoft = ecutdens if ecutsoft is None else ecutsoft
    ecutc = ecut if ecutc is None else ecutc
    ecutorth = ecut if ecutorth is None else ecutorth
    a = _f(np.asarray(rba, np.float64).T.ravel())
    cap = lib().qxo_strvec_capacity(a, float(ecutdens))
    counts = np.zeros(6, np.int32)
    nga, ngb, ngc = (np.zeros(cap, np.int32) for _ in range(3))
    gx, gy, gz, rec = (np.zeros(cap) for _ in range(4))
    cross = np.zeros(9)
    rvol = _d()
    num = lib().qxo_strvec(a, float(ecutdens), float(ecut), float(ecutsoft), float(ecutc),
                           float(ecutorth), cap, counts, nga, ngb, ngc, gx, gy, gz, rec, cross,
                           C.byref(rvol))
    assert num > 0
    return dict(num=int(counts[0]), nplw5=int(counts[1]), nplw=int(counts[2]),
                nplw2=int(counts[3]), nplwc=int(counts[4]), nplw7=int(counts[5]),
                nga=nga, ngb=ngb, ngc=ngc, gx=gx, gy=gy, gz=gz, recnrm=rec,
                cross=cross.reshape(3, 3).T, rvol=rvol.value)


def ijkgd(nga, ngb, ngc, nplw5, nd):
    out = np.zeros(2 * nplw5 + 1, np.int32)
    lib().qxo_ijkgd(int(nplw5), _f(nga, np.int32), _f(ngb, np.int32), _f(ngc, np.int32),
                    _f(nd, np.int32), out)
    return out


def dkgnrm(recnrm, nplwc):
    out = np.zeros(2 * (nplwc + 1))
    lib().qxo_dkgnrm(int(nplwc), _f(recnrm), out)
    return out


def rnd00_sequence(seed, count):
    s = _d(float(seed))
    return np.array([lib().qxo_rnd00(C.byref(s)) for _ in range(count)]), s.value


# ---------------------------------------------------------------- FFT

def rfft3(key, fft3x, fft3y, nd):
    x, y = _f(fft3x).copy(), _f(fft3y).copy()
    lib().qxo_rfft3(int(key), x, y, *[int(v) for v in nd])
    return x, y


def fft3(key, fft3x, fft3y, nd):
    x, y = _f(fft3x).copy(), _f(fft3y).copy()
    lib().qxo_fft3(int(key), x, y, *[int(v) for v in nd])
    return x, y


def _cptr(arr, offset=0):
    return C.c_void_p(arr.ctypes.data + offset * arr.itemsize)


def wvg2r(cgjr, nplw, ijk, nd):
    N = int(np.prod(nd))
    x, y = np.zeros(N), np.zeros(N)
    lib().qxo_wvg2r(_f(cgjr), int(nplw), _cptr(ijk, (len(ijk) - 1) // 2), *[int(v) for v in nd],
                    x, y)
    return x


def hcnonloc(glocal, cgjr, rhcr, nplw, ijk, nd):
    N = int(np.prod(nd))
    r = _f(rhcr).copy()
    x, y = np.zeros(N), np.zeros(N)
    lib().qxo_hcnonloc(_f(glocal), _f(cgjr), r, int(nplw), _cptr(ijk, (len(ijk) - 1) // 2),
                       *[int(v) for v in nd], x, y)
    return r


def ekinetic(dkg, cgjr, occ=None):
    """cgjr: (nb, nplwex).  Returns (rhcr, ekib, ekint)."""
    c = _f(cgjr)
    nb, n = c.shape
    r = np.zeros_like(c)
    ekib = np.zeros(nb)
    ekint = np.zeros(1)
    occ_a = _f(occ) if occ is not None else np.zeros(nb)
    lib().qxo_ekinetic(_f(dkg), c, r, n, nb, _cptr(occ_a), _cptr(ekib), _cptr(ekint))
    return r, ekib, float(ekint[0])


def hckbpp_g(eqlmr, lmx_atom, vclalp_atom, cgjr, rhcr):
    """eqlmr: (natom, lmmax, nplwex); vclalp_atom: (natom, lmmax); cgjr/rhcr: (nb, nplwex).
    Returns (rhcr + KB, slmir[nb, natom, lmmax])."""
    e = _f(eqlmr)
    natom, lmmax, n = e.shape
    c = _f(cgjr)
    nb = c.shape[0]
    r = _f(rhcr).copy()
    sl = np.zeros((nb, natom, lmmax))
    lib().qxo_hckbpp_g(e, n, lmmax, natom, _f(lmx_atom, np.int32), _f(vclalp_atom), c, r, nb,
                       _cptr(sl))
    return r, sl


def xclda(rho, rhocore=None, rdelv=1.0):
    """LDA exchange-correlation (xclda, Sources/vxc.f90:961-1017).  Returns
    vexc, eeexc (Ry) and (eexc, evexc, ecor, evcor)."""
    r = _f(rho)
    n = r.size
    core = None if rhocore is None else _f(rhocore)
    v = np.zeros(n)
    e = np.zeros(n)
    e4 = np.zeros(4)
    lib().qxo_xclda(r, _cptr(core) if core is not None else None, n, float(rdelv), v, e, e4)
    return v, e, e4


def harting(rhgr, recnrmex, ijk, nd):
    """Hartree potential (harting, Sources/engrad.f90:1673-1743): rhgr over G index
    0..nplw5 (real-packed), recnrmex over 0..nplw5, ijk = ijkgd(-nplw5:nplw5).
    Returns vhar on the FFT grid (FFT order, x fastest)."""
    rh = _f(rhgr)
    nplw5 = rh.size // 2 - 1
    ij = _f(ijk, np.int32)
    out = np.zeros(int(np.prod(nd)))
    lib().qxo_harting(rh, nplw5, _f(recnrmex), ij.ctypes.data + 4 * nplw5, int(nd[0]), int(nd[1]),
                      int(nd[2]), out)
    return out


def chgr2g(rho, nplw5, ijk, nd):
    """chgr2g (Sources/engrad.f90:1253-1297): rho (FFT order) -> rhgr over G index
    0..nplw5 (real-packed); ijk = ijkgd(-m:m) with m >= nplw5."""
    ij = _f(ijk, np.int32)
    m = (ij.size - 1) // 2
    out = np.zeros(2 * (int(nplw5) + 1))
    lib().qxo_chgr2g(_f(rho), int(nplw5), ij.ctypes.data + 4 * m, int(nd[0]), int(nd[1]),
                     int(nd[2]), out)
    return out


def chgg2r(rhgr, ijk, nd):
    """chgg2r (Sources/engrad.f90:1302-1353): rhgr over G index 0..nplw5 -> rho
    (FFT order); ijk = ijkgd(-m:m) with m >= nplw5."""
    rh = _f(rhgr)
    ij = _f(ijk, np.int32)
    m = (ij.size - 1) // 2
    out = np.zeros(int(np.prod(nd)))
    lib().qxo_chgg2r(rh, rh.size // 2 - 1, ij.ctypes.data + 4 * m, int(nd[0]), int(nd[1]),
                     int(nd[2]), out)
    return out


def invers(w):
    """INVERS (Sources/funcs.f90:5-61); returns (inverse, ierr)."""
    a = np.asfortranarray(np.asarray(w, np.float64))
    n = a.shape[0]
    flat = np.ascontiguousarray(a.T).ravel()  # column-major as a flat C array
    wi = np.zeros(n * n)
    ierr = lib().qxo_invers(flat, wi, n, n)
    return wi.reshape(n, n).T.copy(), ierr


def factq1(cmetric, g2min, g2max):
    """Metric parameter of the Pulay mixer (Sources/chgdns.f90:661-667)."""
    bbb = g2max - cmetric * g2min
    return (cmetric - 1.0) * g2max * g2min / bbb if bbb > 1.0e-5 else 0.0


class PulayMixer:
    """State of module mxchgg_variables (Sources/chgdns.f90:494-515) for the
    Pulay mixer (imxchg = 1), advanced by mxchgg_pulay's mixing part."""

    def __init__(self, nplwcex, itratn, dkgnrm, rvol, factq1, wkerker=1.0):
        self.nplwcex, self.itratn = int(nplwcex), int(itratn)
        self.dkgnrm = _f(np.asarray(dkgnrm)[: self.nplwcex])
        self.rvol, self.factq1, self.wkerker = float(rvol), float(factq1), float(wkerker)
        self.kntchg = 0
        self.unr = np.zeros(self.itratn * self.nplwcex)
        self.dfmr = np.zeros(self.itratn * self.nplwcex)
        self.aij = np.zeros(self.itratn * self.itratn)
        self.cmk = np.zeros(self.itratn)
        self.nreset = 0

    def mix(self, rhgr, rinr, aslh, bslh, lreset=False, lmixhold=False):
        """Returns (rhgr mixed, rinr) as new arrays; advances the history."""
        rh, ri = _f(rhgr).copy(), _f(rinr).copy()
        k = _i(self.kntchg)
        self.nreset = lib().qxo_mxchgg_pulay(
            rh, ri, self.nplwcex, rh.size, self.dkgnrm, self.rvol, float(aslh), float(bslh),
            self.factq1, self.wkerker, int(bool(lreset)), int(bool(lmixhold)), self.itratn,
            C.byref(k), self.unr, self.dfmr, self.aij, self.cmk)
        self.kntchg = k.value
        return rh, ri


def mxchgg_simple(rhgr, rinr, aslh, none=False):
    """mxchgg_simple / mxchgg_none mixing part (Sources/chgdns.f90:907-1210)."""
    rh, ri = _f(rhgr).copy(), _f(rinr).copy()
    lib().qxo_mxchgg_simple(rh, ri, rh.size, float(aslh), int(bool(none)))
    return rh, ri


def axis_tables(frac, kmax):
    """DCQ/DSQ of stscfn_pw for periodic cells: cos/sin(2 pi k s_a) for
    k = -kmax_a..kmax_a, one (natom, 2 kmax_a + 1) pair per axis a."""
    fr = np.asarray(frac, np.float64)
    out = []
    for a in range(3):
        k = np.arange(-kmax[a], kmax[a] + 1)
        ph = 2 * np.pi * np.outer(fr[:, a], k)
        out.append((np.ascontiguousarray(np.cos(ph)), np.ascontiguousarray(np.sin(ph))))
    return out


def flocal_g(ntype, nhk1, nhk2, iatmpt, llking, natom, nga, ngb, ngc, gx, gy, gz, rhgr, svco,
             kmax, tables, floc=None):
    """flocal_g without stress (Sources/nlkbpp.f90:1037-1125).  svco (ntype,
    nplw5 + 1) over ig = 0..nplw5; tables from axis_tables; returns floc (natom, 3)."""
    rh = _f(rhgr)
    n5 = rh.size // 2 - 1
    f = np.zeros((natom, 3)) if floc is None else _f(floc).copy()
    sv = _f(svco)
    dc = (_vp * 3)(*[t[0].ctypes.data for t in tables])
    ds = (_vp * 3)(*[t[1].ctypes.data for t in tables])
    lib().qxo_flocal_g(int(ntype), _f(nhk1, np.int32), _f(nhk2, np.int32), _f(iatmpt, np.int32),
                       _f(llking, np.int32), int(natom), n5, _f(np.asarray(nga)[: n5 + 1], np.int32),
                       _f(np.asarray(ngb)[: n5 + 1], np.int32), _f(np.asarray(ngc)[: n5 + 1], np.int32),
                       _f(np.asarray(gx)[: n5 + 1]), _f(np.asarray(gy)[: n5 + 1]),
                       _f(np.asarray(gz)[: n5 + 1]), rh, sv, _f(kmax, np.int32), dc, ds, f)
    return f


def vlocal_g(ntype, nhk1, nhk2, iatmpt, llking, natom, nga, ngb, ngc, svco, rvol, kmax, tables,
             ijk, nd):
    """vlocal_g (Sources/nlkbpp.f90:862-1032) for a zeroed vext: returns (vext
    on the grid in FFT order, thrhgr over ig = 0..nplw5).  The G = 0 grid
    value and the +-G scatter are those of chgg2r's inverse transform."""
    sv = _f(svco)
    n5 = sv.shape[-1] - 1
    thr = np.zeros(2 * (n5 + 1))
    dc = (_vp * 3)(*[t[0].ctypes.data for t in tables])
    ds = (_vp * 3)(*[t[1].ctypes.data for t in tables])
    v0 = lib().qxo_vlocal_g(int(ntype), _f(nhk1, np.int32), _f(nhk2, np.int32),
                            _f(iatmpt, np.int32), _f(llking, np.int32), int(natom), n5,
                            _f(np.asarray(nga)[: n5 + 1], np.int32),
                            _f(np.asarray(ngb)[: n5 + 1], np.int32),
                            _f(np.asarray(ngc)[: n5 + 1], np.int32), sv, float(rvol),
                            _f(kmax, np.int32), dc, ds, thr)
    z = thr.copy()
    z[0], z[1] = v0, 0.0
    return chgg2r(z, ijk, nd), thr


def frcrec(ntype, nhk1, nhk2, zv, nel, natom, iatoit, iatmpt, ew, tables, fclm=None,
           ecolmb=0.0, atom_ecolmb=None):
    """frcrec without stress (Sources/bulk.f90:1880-2026).  ew: the Ewald G
    list dict (nga, ngb, ngc, gx, gy, gz, recnrm over m = 1..numk2 at [m-1],
    recnrmex over 0..numk2, gamma, volume); tables: per-ion axis tables
    [(DCQ_a, DSQ_a)] (nion, 2 kcsma_a + 1).  Returns (fclm (natom, 3),
    ecolmb, atom_ecolmb)."""
    nion = len(iatmpt)
    f = np.zeros((natom, 3)) if fclm is None else _f(fclm).copy()
    ae = np.zeros(natom) if atom_ecolmb is None else _f(atom_ecolmb).copy()
    e = _d(ecolmb)
    km = np.array([(t[0].shape[1] - 1) // 2 for t in tables], np.int32)
    dc = (_vp * 3)(*[t[0].ctypes.data for t in tables])
    ds = (_vp * 3)(*[t[1].ctypes.data for t in tables])
    n = len(ew["nga"])
    lib().qxo_frcrec(int(ntype), _f(nhk1, np.int32), _f(nhk2, np.int32), _f(zv), int(nel),
                     int(natom), nion, _f(iatoit, np.int32), _f(iatmpt, np.int32), n,
                     _f(ew["nga"], np.int32), _f(ew["ngb"], np.int32), _f(ew["ngc"], np.int32),
                     _f(ew["gx"]), _f(ew["gy"]), _f(ew["gz"]), _f(ew["recnrm"]),
                     _f(ew["recnrmex"]), float(ew["gamma"]), float(ew["volume"]), km, dc, ds, f,
                     C.byref(e), ae)
    return f, e.value, ae


def roteig(cur, prev, kpwvx):
    """roteig2 (Sources/sbalin.f90:772-1373): prev (nprvmx, nb, nplwex) is
    rotated in place (sets 1..min(nprvmx, kpwvx)); returns (prev, U, rc)."""
    c = _f(cur)
    p = _f(prev).copy()
    nprv, nb, n = p.shape
    u = np.zeros(nb * nb)
    rc = lib().qxo_roteig(c, p, n, nb, nprv, int(kpwvx), u.ctypes.data)
    return p, u.reshape(nb, nb).T.copy(), rc


def coef_aspc(kaspc):
    """set_coef_ASPC's coefaspc (kaspc, kaspc) as coef[j-1, k-1]."""
    c = np.zeros(kaspc * kaspc)
    lib().qxo_coef_aspc(int(kaspc), c)
    return c.reshape(kaspc, kaspc).T.copy()


class Aspc:
    """sbalinaspc state (Sources/sbalin.f90:369-712): the history prveig of
    nprvmx sets and keypwv; step() advances one MD step."""

    def __init__(self, nb, nplwex, nprvmx, ihest):
        self.nprvmx, self.ihest = int(nprvmx), int(ihest)
        self.kaspc = max(self.ihest, 0) + 1
        self.coef = _f(np.ascontiguousarray(coef_aspc(self.kaspc).T).ravel())
        self.prveig = np.zeros((self.nprvmx, nb, nplwex))
        self.keypwv = 0

    def step(self, gdcr, nstep, nstep_ini=0):
        g = _f(gdcr).copy()
        nb, n = g.shape
        k = _i(self.keypwv)
        ia = lib().qxo_sbalinaspc(g, n, nb, self.prveig, self.nprvmx, self.coef, self.kaspc,
                                  self.ihest, int(nstep), int(nstep_ini), C.byref(k))
        self.keypwv = k.value
        return g, ia


def kbnlpp_frc2(eqlmr, lmx_atom, vclalp_atom, cgjr, occ, slmir, gx, gy, gz):
    """KB Hellmann-Feynman forces (kbnlpp_frc2, Sources/nlkbpp.f90:1878-2060):
    eqlmr (natom, lmmax, nplwex), cgjr (nb, nplwex), slmir (nb, natom, lmmax) from
    hckbpp_g, gx/gy/gz over G index 0..nplw.  Returns fnlc (natom, 3)."""
    e = _f(eqlmr)
    natom, lmmax, n = e.shape
    c = _f(cgjr)
    f = np.zeros((natom, 3))
    nw = n // 2
    lib().qxo_kbnlpp_frc2(e, n, lmmax, natom, _f(lmx_atom, np.int32), _f(vclalp_atom), c,
                          c.shape[0], _f(occ), _f(slmir), _f(np.asarray(gx)[:nw]),
                          _f(np.asarray(gy)[:nw]), _f(np.asarray(gz)[:nw]), f)
    return f


def stscfn_pw(ratm, hcell, nd1v, nd1vks, nga, ngb, ngc, nplwcs, kmax, lclust=False):
    """Structure factors (stscfn_pw, Sources/pwlda.f90:5711-5854).  ratm (natom, 3)
    scaled coordinates, hcell (3, 3) with lattice vectors as columns, nga/ngb/ngc
    for G index 0..nplw5.  Returns ycos, ysin of shape (natom, nplwcs + 1)."""
    r = _f(ratm)
    natom = r.shape[0]
    ga = _f(nga, np.int32)
    yc = np.zeros((natom, nplwcs + 1))
    ys = np.zeros((natom, nplwcs + 1))
    lib().qxo_stscfn_pw(r, _f(np.asarray(hcell).T), _f(nd1v, np.int32), _f(nd1vks, np.int32),
                        int(bool(lclust)), natom, int(kmax[0]), int(kmax[1]), int(kmax[2]),
                        int(nplwcs), len(ga) - 1, ga, _f(ngb, np.int32), _f(ngc, np.int32), yc, ys)
    return yc, ys


def setnlc_g(t, nhk1, nhk2, natom, ycos, ysin, recnrm, rvol, lkbppi=None, lking=None):
    """KB table (setnlc_g / setbgq / setqlm, Sources/nlkbpp.f90:1236-1279, 2135-2247).
    t: nlpp tables (qxmd_amd.nlpp.NlppTables); ycos/ysin (natom, nplwcs + 1);
    recnrm (nplw + 1).  Returns eqlmr (natom, lmmax, nplwex), vclalp (ntype, lmmax)."""
    nplw = len(recnrm) - 1
    ntype, lmmax = t.itola.shape
    yc = _f(ycos)
    nplwcs = yc.shape[1] - 1
    lk = np.ones(ntype, np.int32) if lkbppi is None else _f(lkbppi, np.int32)
    lg = np.zeros(ntype, np.int32) if lking is None else _f(lking, np.int32)
    eq = np.zeros((natom, lmmax, 2 * (nplw + 1)))
    vcl = np.zeros((ntype, lmmax))
    lib().qxo_setnlc_g(nplw, ntype, _f(nhk1, np.int32), _f(nhk2, np.int32), natom, lk, lg, yc,
                       _f(ysin), nplwcs, float(rvol), lmmax, t.tbqlm.shape[1],
                       t.tbqlm.shape[2] - 1, _f(t.lmx, np.int32), _f(t.clalp), _f(t.tbqlm),
                       _f(t.tbqlma), _f(t.dlqlm), _f(t.itola, np.int32), _f(t.itolm, np.int32),
                       _f(t.itolmf, np.int32), _f(t.litoll, np.int32), t.qlmr.shape[0],
                       _f(t.qlmr[:, :nplw + 1]), _f(t.qlmi[:, :nplw + 1]), _f(recnrm), eq, vcl)
    return eq, vcl


def enkbpp_g(slmir, lmx_atom, clalp_atom, occ, rvol):
    sl = _f(slmir)
    nb, natom, lmmax = sl.shape
    return lib().qxo_enkbpp_g(sl, lmmax, natom, _f(lmx_atom, np.int32), _f(clalp_atom), _f(occ),
                              nb, float(rvol))


def rsdecmp_rsinvcd(s):
    """Cholesky + inverse of the lower factor; s symmetric (n, n).  Returns Linv (lower)."""
    n = s.shape[0]
    a = _f(np.asarray(s).T).copy()  # column-major
    lib().qxo_rsdecmp(a, n, n)
    w = np.zeros(n)
    lib().qxo_rsinvcd(a, n, n, w)
    return a.reshape(n, n).T


def schmidt(cgjr):
    c = _f(cgjr).copy()
    nb, n = c.shape
    s = np.zeros(nb * nb)
    ok = lib().qxo_schmidt(c, n, nb, s)
    return c, bool(ok)


def hmatrix(cgjr, rhcr):
    c, r = _f(cgjr), _f(rhcr)
    nb, n = c.shape
    h = np.zeros(nb * nb)
    eig = np.zeros(nb)
    lib().qxo_hmatrix(c, r, n, nb, h, eig)
    return h.reshape(nb, nb).T.copy(), eig


def eigh(a):
    n = a.shape[0]
    m = _f(np.asarray(a).T).copy()
    w = np.zeros(n)
    lib().qxo_eigh(m, n, n, w)
    return w, m.reshape(n, n).T.copy()


def untrycore(gdcr, u):
    g = _f(gdcr).copy()
    nb, n = g.shape
    work = np.zeros_like(g)
    lib().qxo_untrycore(g, n, nb, _f(np.asarray(u).T), nb, work)
    return g


def softchg(cgjr, nplw, occ, rvol, ijk, nd):
    c = _f(cgjr)
    nb, n = c.shape
    rho = np.zeros(int(np.prod(nd)))
    lib().qxo_softchg(c, n, int(nplw), nb, _f(occ), float(rvol), _cptr(ijk, (len(ijk) - 1) // 2),
                      *[int(v) for v in nd], rho)
    return rho


def prpcg(ekib, dkg):
    out = np.zeros(len(dkg))
    lib().qxo_prpcg(float(ekib), _f(dkg), len(dkg), out)
    return out


# ---------------------------------------------------------------- Hamiltonian

class Hamiltonian:
    """Holds the arrays a ``Ham`` struct points to (keeps them alive)."""

    def __init__(self, nplw, nd, ijk, glocal, dkg, eqlmr=None, lmx_atom=None, vclalp_atom=None):
        self.nplw = int(nplw)
        self.nplwex = 2 * (self.nplw + 1)
        self.nd = tuple(int(v) for v in nd)
        self.ijk = _f(ijk, np.int32)
        self.glocal = _f(glocal)
        self.dkg = _f(dkg)[: self.nplwex].copy()
        self.eqlmr = None if eqlmr is None else _f(eqlmr)
        self.lmx = None if lmx_atom is None else _f(lmx_atom, np.int32)
        self.vcl = None if vclalp_atom is None else _f(vclalp_atom)
        h = Ham()
        h.nplw, h.nplwex = self.nplw, self.nplwex
        h.n1, h.n2, h.n3 = self.nd
        h.ijkgd_c = _cptr(self.ijk, (len(self.ijk) - 1) // 2)
        h.glocal = _cptr(self.glocal)
        h.dkgnrm = _cptr(self.dkg)
        if self.eqlmr is not None:
            h.eqlmr = _cptr(self.eqlmr)
            h.natom, h.lmmax = self.eqlmr.shape[0], self.eqlmr.shape[1]
            h.lmx_atom = _cptr(self.lmx)
            h.vclalp_atom = _cptr(self.vcl)
        else:
            h.eqlmr = None
            h.natom = h.lmmax = 0
        self.h = h

    def hpsi(self, psi):
        p = _f(psi)
        out = np.zeros_like(p)
        lib().qxo_hpsi(C.byref(self.h), p, out, p.shape[0], None)
        return out

    def eigen_bkl(self, eig, gdcr, rhcr, ekib, noband=0, itermx=4, iteremx=4, tolera=1e-10,
                  threinn=0.0, wdinner=0.9, roughfzan=0.3, roughfeig=0.3):
        e, g, r = _f(eig).copy(), _f(gdcr).copy(), _f(rhcr).copy()
        nb = g.shape[0]
        bzan = np.zeros(nb)
        nriter = np.zeros(nb, np.int32)
        lib().qxo_eigen_bkl(C.byref(self.h), e, g, r, _f(ekib), nb, int(noband), int(itermx),
                            int(iteremx), tolera, threinn, wdinner, roughfzan, roughfeig, bzan,
                            nriter)
        return e, g, r, bzan, nriter

    BKL_TRACE_FIELDS = ("ib", "iter", "h11", "zansap", "beta", "cai", "cbi", "hhi", "eig",
                        "dtnew", "alpsq", "zansa", "bkl", "stop", "eigb", "z000re", "x000re")

    def eigen_bkl_trace(self, eig, gdcr, rhcr, ekib, cap=4096, vectors=True, **kw):
        """eigen_bkl with its per-iteration trace: returns (eig, gdcr, rhcr,
        bzan, nriter, records, vecs) where records is a structured list of
        dicts (BKL_TRACE_FIELDS) and vecs[i] = (x before the update, h after
        the projections) of record i."""
        n = _f(gdcr).shape[1]
        rec = np.zeros((cap, len(self.BKL_TRACE_FIELDS)))
        vec = np.zeros((cap, 2, n)) if vectors else None
        lib().qxo_bkl_trace(rec.ctypes.data, vec.ctypes.data if vectors else None, cap)
        try:
            out = self.eigen_bkl(eig, gdcr, rhcr, ekib, **kw)
            cnt = lib().qxo_bkl_trace_count()
        finally:
            lib().qxo_bkl_trace(None, None, 0)
        recs = [dict(zip(self.BKL_TRACE_FIELDS, r)) for r in rec[:cnt]]
        for r in recs:
            for k in ("ib", "iter", "bkl", "stop"):
                r[k] = int(r[k])
        return (*out, recs, vec[:cnt] if vectors else None)

    def scf_band_step(self, gdcr, occ, rvol, noband=0, clalp_atom=None, want_rho=True):
        g = _f(gdcr).copy()
        nb = g.shape[0]
        r = np.zeros_like(g)
        eig = np.zeros(nb)
        rho = np.zeros(int(np.prod(self.nd))) if want_rho else None
        ek, en = _d(), _d()
        cl = _f(clalp_atom) if clalp_atom is not None else None
        ok = lib().qxo_scf_band_step(C.byref(self.h), g, r, eig, _f(occ), nb, int(noband),
                                     float(rvol), _cptr(rho) if rho is not None else None,
                                     C.byref(ek), C.byref(en), _cptr(cl) if cl is not None else None)
        return dict(psi=g, hpsi=r, eig=eig, rho=rho, ekin=ek.value, enl=en.value, ok=bool(ok))


# ---- velocity-Verlet updates of the MD loop at constant energy -------------
# Scaled coordinates: x in cell units, v the displacement per step, forces
# turned into a = acon H^-1 f (acon = dt^2 / 2m per species,
# Sources/remd.f90:3568-3571).  Scalar Python floats, in the reference's
# operand order.

def rciprl(hcell):
    """Sources/reclat.f90:1089-1128: (b, v), b[:, j] the j-th reciprocal
    vector / 2 pi of the cell with columns hcell[:, j], v = |det|."""
    a = lambda i, j: float(hcell[i - 1][j - 1])  # noqa: E731
    b = [[0.0] * 3 for _ in range(3)]
    b[0][0] = a(2, 2) * a(3, 3) - a(3, 2) * a(2, 3)
    b[1][0] = a(3, 2) * a(1, 3) - a(1, 2) * a(3, 3)
    b[2][0] = a(1, 2) * a(2, 3) - a(2, 2) * a(1, 3)
    v = a(1, 1) * b[0][0] + a(2, 1) * b[1][0] + a(3, 1) * b[2][0]
    vr = 1.0 / v
    for i in range(3):
        b[i][0] = b[i][0] * vr
    b[0][1] = (a(2, 3) * a(3, 1) - a(3, 3) * a(2, 1)) * vr
    b[1][1] = (a(3, 3) * a(1, 1) - a(1, 3) * a(3, 1)) * vr
    b[2][1] = (a(1, 3) * a(2, 1) - a(2, 3) * a(1, 1)) * vr
    b[0][2] = (a(2, 1) * a(3, 2) - a(3, 1) * a(2, 2)) * vr
    b[1][2] = (a(3, 1) * a(1, 2) - a(1, 1) * a(3, 2)) * vr
    b[2][2] = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) * vr
    return np.array(b), abs(v)


def normalized_a(frc, species, hcell, acon):
    """Sources/remd.f90:6789-6830: a_i = acon(is_i) (hi(:,k) . f_i)_k."""
    hi, _ = rciprl(hcell)
    out = np.array(frc, dtype=np.float64, copy=True)
    for i, f in enumerate(out):
        arx, ary, arz = float(f[0]), float(f[1]), float(f[2])
        ax = hi[0, 0] * arx + hi[1, 0] * ary + hi[2, 0] * arz
        ay = hi[0, 1] * arx + hi[1, 1] * ary + hi[2, 1] * arz
        az = hi[0, 2] * arx + hi[1, 2] * ary + hi[2, 2] * arz
        c = float(acon[species[i] - 1])
        out[i] = (ax * c, ay * c, az * c)
    return out


def vkick(v, a, species, lfixion):
    """Sources/remd.f90:5408-5430: v + a for atoms of unfixed species
    (updtps, 5755-5783, is the same update of x by v)."""
    out = np.array(v, dtype=np.float64, copy=True)
    for i in range(out.shape[0]):
        if not lfixion[species[i] - 1]:
            out[i] = out[i] + np.asarray(a[i], np.float64)
    return out


updtps = vkick


def updtps_pbc(x):
    """Sources/remd.f90:5786-5806: scaled coordinates into [0, 1)."""
    out = np.array(x, dtype=np.float64, copy=True)
    flat = out.reshape(-1)
    for k in range(flat.size):
        xi = float(flat[k])
        if xi < 0.0:
            xi = xi - float(int(xi)) + 1.0
        if xi >= 1.0:
            xi = xi - float(int(xi))
        flat[k] = xi
    return out
