"""The C-ABI library loads and exports every entry point include/autoplace_gp.h declares (no GPU call)."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "autoplace_gp.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"AP_API\s+[\w\s*]*?\b(ap_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("ap_wa_wirelength", "ap_density_map", "ap_nesterov_step", "ap_density_grad", "ap_control",
              "ap_select_step", "ap_poisson_solve", "ap_level_run"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from autoplace_amd import _lib
    lib = _lib.lib()
    for n in _declared():
        assert hasattr(lib, n), n
        ctypes.cast(getattr(lib, n), ctypes.c_void_p)
    assert set(_lib.exported_symbols()) == set(_declared())
    assert lib.ap_version() >= 100


def test_argument_errors_are_reported_without_a_gpu():
    """Argument checks run on the host before any launch."""
    import pytest
    from autoplace_amd import _lib
    lib = _lib.lib()
    st = lib.ap_density_map(None, None, None, 0, 0, 1.0, 1.0, 0, 4, None, None, 1.0, None, None, None)
    assert st == _lib.AP_ERR_ARG
    assert b"bad grid" in lib.ap_last_error()
    with pytest.raises(ValueError):
        _lib.check(st, "density_map")
