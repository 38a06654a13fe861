"""ctypes binding of libqxmd_amd.so (the C-ABI declared in include/qxmd_amd.h).

The library is the product: nothing here falls back to a CPU path.  Importing
works without a GPU (for set-up calls and symbol checks); any compute call on
a machine without a HIP device fails with QxError.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "libqxmd_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "qxmd_amd.h")

_LIB = None


class QxError(RuntimeError):
    pass


_LIVE = weakref.WeakSet()  # open device contexts (PlaneWave), closed at exit


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(SO):
            raise QxError(f"{SO} is missing: build it with `python -m qxmd_amd.build`")
        _LIB = C.CDLL(SO)
        _declare(_LIB)
        atexit.register(_teardown)
    return _LIB


def track(ctx):
    """Register an object with a close() method to be closed at interpreter
    exit, before the HIP runtime's own teardown."""
    _LIVE.add(ctx)


def _teardown():
    """Release every open context and the drop-ins' process-wide device
    resources while the HIP runtime is still up."""
    for ctx in list(_LIVE):
        try:
            ctx.close()
        except Exception:
            pass
    if _LIB is not None:
        _LIB.qx_shutdown()


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i, _d, _vp = C.c_int, C.c_double, C.c_void_p
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)

# name -> argtypes (restype int unless noted)
SIGNATURES = {
    "qx_version": [],
    "qx_get_mesh": [_dp, _d, _ip],
    "qx_strvec": [_dp, _d, _d, _i, _ip, _ip, _ip, _ip, _dp, _vp, _vp, _vp],
    "qx_ijkgd": [_i, _ip, _ip, _ip, _ip, _ip],
    "qx_clmrcp": [_dp, _d, _i, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _ip],
    "qx_create": [C.POINTER(_vp), _i, _i, _i, _i, _ip, _ip, _ip, _dp, _i],
    "qx_destroy": [_vp],
    "qx_sync": [_vp],
    "qx_shutdown": [],
    "qx_set_glocal": [_vp, _dp],
    "qx_set_kb": [_vp, _dp, _i, _i, _ip, _dp],
    "qx_info": [_vp, _ip],
    "qx_set_nlpp_tables": [_vp, _i, _i, _i, _i, _ip, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _i,
                           _i, _dp, _dp],
    "qx_stscfn_pw": [_vp, _dp, _i, _dp, _ip, _ip, _i, _i, _i, _i, _i, _i, _ip, _ip, _ip, _vp,
                     _vp],
    "qx_setnlc_g": [_vp, _i, _i, _ip, _ip, _i, _ip, _ip, _vp, _vp, _i, _d, _vp],
    "qx_get_kb": [_vp, _dp, _vp],
    "qx_kbnlpp_frc2_slot": [_vp, _i, _i, _dp, _dp, _dp, _dp, _dp],
    "qx_kbnlpp_frc2": [_vp, _dp, _i, _dp, _dp, _dp, _dp, _dp],
    "qx_slot_put": [_vp, _i, _dp, _i],
    "qx_slot_get": [_vp, _i, _dp, _i],
    "qx_slot_copy": [_vp, _i, _i, _i],
    "qx_hpsi_slot": [_vp, _i, _i, _i, _i],
    "qx_local_slot": [_vp, _i, _i, _i, _i, _i],
    "qx_kb_slot": [_vp, _i, _i, _i, _vp],
    "qx_schmidt_slot": [_vp, _i, _i, _pi],
    "qx_rotate_slot": [_vp, _i, _i, _i, _dp],
    "qx_density_slot": [_vp, _i, _i, _dp, _d, _vp, _i],
    "qx_eigen_bkl_slot": [_vp, _i, _i, _i, _dp, _dp, _i, _i, _i, _d, _d, _d, _d, _d, _dp, _ip, _i],
    "qx_scf_band_step": [_vp, _i, _i, _i, _dp, _i, _d, _dp, _vp, _pi, _i],
    "qx_event_record": [_vp, _i],
    "qx_event_elapsed": [_vp, _i, _i, C.POINTER(C.c_float)],
    "qx_scf_phase_ms": [_vp, C.POINTER(C.c_float)],
    "qx_launch_timer": [_vp, _i, _i, C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "qx_rfft3": [_i, _dp, _dp, _i, _i, _i, _pi],
    "qx_fft3": [_i, _dp, _dp, _i, _i, _i, _pi],
    "qx_rfft3_dev_noalloc": [_i, _vp, _vp, _i, _i, _i, _vp, _pi],
    "qx_noalloc_limits": [C.POINTER(C.c_long), _pi],
    "qx_set_gemm_path": [_i],
    "qx_set_cg_graph": [_vp, _i],
    "qx_cg_graph_count": [_vp],
    "qx_gemm_path": [],
    "qx_set_gemm_schedule": [_i],
    "qx_gemm_schedule": [],
    "qx_hcnonloc": [_vp, _dp, _dp],
    "qx_hckbpp_g": [_vp, _dp, _dp, _i, _vp],
    "qx_hckbpp_ib_g": [_vp, _dp, _dp],
    "qx_schmidt": [_vp, _dp, _i, _pi],
    "qx_softchg": [_vp, _dp, _i, _dp, _d, _dp],
    "qx_rseigqr": [_vp, _vp, _i, _dp],
    "qx_icholesky": [_vp, _vp, _i, _pi],
    "qx_set_lopp": [_vp, _i, _i, _dp, _dp, _dp, _dp],
    "qx_flocal_g": [_vp, _dp, _i, _ip, _ip, _ip, _i, _ip, _i, _ip, _ip, _ip, _dp, _i, _i, _i,
                    _dp, _dp, _dp, _dp, _dp, _dp],
    "qx_flocal_g_dev": [_vp, _i, _ip, _ip, _ip, _ip, _i, _dp],
    "qx_vlocal_g": [_vp, _i, _ip, _ip, _ip, _ip, _i, _d, _i, _vp, _vp],
    "qx_svwfns": [C.c_char_p, _dp, _i, _i, _i, _dp, _dp, _i],
    "qx_rdwfns": [C.c_char_p, _dp, _i, _i, _i, _dp, _dp, _pi],
    "qx_svwfns_slot": [_vp, C.c_char_p, _i, _i, _dp, _dp, _i],
    "qx_rdwfns_slot": [_vp, C.c_char_p, _i, _i, _dp, _dp, _pi],
    "qx_svacon": [C.c_char_p, _i, _ip, _ip, _i, _i, _i, _dp, _i, _dp, _dp, _dp],
    "qx_rdacon": [C.c_char_p, _i, _ip, _ip, _pi, _pi, _pi, _dp, _pi, _dp, _dp, _dp],
    "qx_rciprl": [_dp, _dp, _pd],
    "qx_normalized_a": [_i, _i, _dp, _ip, _dp, _dp],
    "qx_vkick": [_dp, _dp, _ip, _i, _i, _ip],
    "qx_updtps": [_dp, _dp, _ip, _i, _i, _ip],
    "qx_updtps_pbc": [_dp, _i],
    "qx_aspc_init": [_vp, _i, _i],
    "qx_sbalinaspc_slot": [_vp, _i, _i, _i, _i, _vp],
    "qx_roteig_slot": [_vp, _i, _i, _i],
    "qx_aspc_get": [_vp, _i, _i, _dp],
    "qx_aspc_put": [_vp, _i, _i, _dp, _i],
    "qx_set_ewald": [_vp, _i, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _d, _d],
    "qx_frcrec": [_vp, _dp, _i, _ip, _ip, _dp, _i, _pd, _dp, _ip, _ip, _i, _i, _i, _i, _i, _dp,
                  _dp, _dp, _dp, _dp, _dp],
    "qx_frcrec_dev": [_vp, _dp, _i, _ip, _ip, _dp, _i, _pd, _dp, _ip, _ip, _i, _i],
    "qx_xclda": [_dp, _vp, _i, _d, _dp, _dp, _dp],
    "qx_harting": [_dp, _dp, _i, _dp, _vp, _i, _ip, _ip],
    "qx_chgr2g": [_dp, _dp, _i, _vp, _i, _ip, _ip],
    "qx_chgg2r": [_dp, _dp, _i, _vp, _i, _ip, _ip],
    "qx_mxchgg_alloc": [_vp, _i, _i, _i, _d, _d, _d, _d, _dp, _d],
    "qx_mxchgg": [_vp, _pi, _dp, _dp, _i, _d, _d, _dp, _vp, _i, _ip, _i],
    "qx_scf_set_grid": [_vp, _i, _ip, _dp, _vp, _vp, _d],
    "qx_pwpchg_slot": [_vp, _i, _i, _dp, _d, _i],
    "qx_outchg": [_vp],
    "qx_pwpmxchgg": [_vp, _i, _d, _d, _i, _vp],
    "qx_scf_potential": [_vp, _dp],
    "qx_scf_get": [_vp, _i, _dp],
    "qx_scf_put": [_vp, _i, _dp],
    "qx_ksphere": [_dp, _dp, _d, _ip, _i, _vp, _vp],
    "qx_kpoints_set": [_vp, _i, _dp, _dp, _d, _i, _ip],
    "qx_kpoints_sphere": [_vp, _i, _ip, _dp],
    "qx_kslot_put": [_vp, _i, _i, _dp, _i],
    "qx_kslot_get": [_vp, _i, _i, _dp, _i],
    "qx_kset_kb": [_vp, _i, _i, _dp, _dp],
    "qx_khpsi_slot": [_vp, _i, _i, _i, _i],
    "qx_ksolve_slot": [_vp, _i, _i, _i, _d, _dp, _ip],
}
# entry points beyond the header's core set (batched CG, kinetic energies)
EXTRA = {
    "qx_eigen_bkl_batched_slot": [_vp, _i, _i, _i, _dp, _dp, _i, _i, _i, _d, _d, _d, _dp, _ip, _i],
    "qx_ekib_slot": [_vp, _i, _i, _dp],
    "qx_scf_band_step_ex": [_vp, _i, _i, _i, _dp, _i, _d, _dp, _vp, _pi, _i, _i],
}


def _declare(L):
    for name, args in {**SIGNATURES, **EXTRA}.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.qx_last_error.argtypes = []
    L.qx_last_error.restype = C.c_char_p


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().qx_last_error().decode(errors="replace")
        raise QxError(f"{what}: qxmd_amd error {rc}: {msg}")


def declared_symbols():
    """Names of the functions include/qxmd_amd.h declares."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qx_[a-z0-9_]+)\s*\(", txt)))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
