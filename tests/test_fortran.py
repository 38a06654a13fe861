"""The Fortran host layer: ISO_C_BINDING interfaces and the drop-in externals
with the reference's names (rfft3, fft3, RsEIGQR, icholesky, softchg,
hckbpp_g, hckbpp_ib_g, stscfn_pw, kbnlpp_frc2, xclda, harting, chgr2g, chgg2r,
mxchgg_alloc, set_g2ming2max_in_chgdns, mxchgg, flocal_g, updtps, vkick,
updtps_pbc, normalized_a).  The CPU test checks the built library exports them;
the GPU test runs the Fortran driver that calls each by its reference
argument list and checks closed forms."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLIB = os.path.join(ROOT, "qxmd_amd", "libqxmd_amd_f.so")
DRIVER = os.path.join(ROOT, "tests", "fortran", "dropin_check")
DROPINS = ["rfft3_", "fft3_", "rseigqr_", "icholesky_", "softchg_", "hckbpp_g_", "hckbpp_ib_g_",
           "stscfn_pw_", "kbnlpp_frc2_", "xclda_", "harting_", "chgr2g_", "chgg2r_", "mxchgg_",
           "mxchgg_alloc_", "set_g2ming2max_in_chgdns_", "flocal_g_", "updtps_", "vkick_",
           "updtps_pbc_", "normalized_a_"]

needs_fortran = pytest.mark.skipif(not os.path.exists(FLIB),
                                   reason="Fortran layer not built (no flang)")


@needs_fortran
def test_fortran_layer_exports_dropins():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", FLIB], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in DROPINS if s not in names]
    assert not missing, missing
    assert os.path.exists(DRIVER)


@pytest.mark.gpu
@needs_fortran
def test_fortran_dropin_driver():
    r = subprocess.run([DRIVER], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300, cwd=ROOT)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "DROPIN OK" in r.stdout


MODULE_DRIVER = os.path.join(ROOT, "tests", "fortran", "dropin_modules_check")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(MODULE_DRIVER),
                    reason="built only where the reference's flang build (oracle/_ref) exists")
def test_fortran_module_dropins():
    """schmidt, pwpchg and eigen by name, on the reference's own module data."""
    r = subprocess.run([MODULE_DRIVER], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300, cwd=ROOT)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "MODULE DROPIN OK" in r.stdout
