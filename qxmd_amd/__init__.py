"""qxmd_amd -- MI355X-native Kohn-Sham H|psi> / band-update path of QXMD.

The compute path is the HIP library ``libqxmd_amd.so`` (sources in ``csrc/``,
C-ABI in ``include/qxmd_amd.h``); this package holds its ctypes binding
(``lib``), the host-side mirror of the reference's plane-wave interface
(``pw``), and synthetic-cell builders for tests and benchmarks (``system``).
"""
from .lib import QxError, load  # noqa: F401
