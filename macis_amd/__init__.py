"""MI355X-native determinant-CI hot path (MACIS Hamiltonian build, SpMV,
Davidson, ASCI search) behind a C-ABI (include/macis_hip.h).

The compute lives in libmacis_hip.so (HIP kernels for gfx950); this package
only binds it (``_capi``) and provides input helpers (FCIDUMP reader,
determinant lists, seeded synthetic integrals).
"""
from . import dets, fcidump, synthetic  # noqa: F401

__all__ = ["dets", "fcidump", "synthetic", "capi"]


def capi():
    """The ctypes binding; raises if libmacis_hip.so is not built."""
    from . import _capi

    _capi.lib()
    return _capi
