"""CPU checks of the drop-in boundary: the HIP library loads and exports
every entry point declared in include/macis_hip.h (no compute calls)."""
import ctypes

from macis_amd import _capi


def test_library_exports_every_declared_symbol():
    names = _capi.declared_symbols()
    assert len(names) >= 25
    L = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_declared_symbols():
    assert set(_capi.declared_symbols()) == set(_capi._SIGS)


def test_version_and_error_strings():
    L = _capi.lib()
    assert L.mhip_get_version() == 10000
    assert L.mhip_error_string(-5).decode() == "Davidson Did Not Converge!"
    assert L.mhip_error_string(-6).decode() == "Davidson: No Guess Provided"
    assert L.mhip_error_string(-7).decode() == "Unsupported index_t"
