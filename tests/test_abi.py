"""The C ABI library loads and exports every symbol include/mgc.h declares (no GPU)."""
import ctypes

import mgc


def test_library_exports_every_declared_symbol():
    names = mgc.declared_symbols()
    assert len(names) > 30
    L = ctypes.CDLL(mgc.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_declared_symbols():
    # every declared entry point has ctypes argument types in the binding
    names = mgc.declared_symbols()
    assert not [n for n in names if n not in mgc._SIGS]


def test_errors_are_status_codes_without_device():
    L = mgc.lib()
    # argument validation happens before any device call
    assert L.mgc_coarse_op_create(None, None, 8, mgc.F64, 0) == -1
    assert b"null" in L.mgc_last_error()
    assert L.mgc_coarse_op_M(None, None, None, None) == -1
    assert L.mgc_version() >= 10000
