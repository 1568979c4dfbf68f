"""CPU checks of the C-ABI library: it loads without a GPU and exports every
entry point include/mfx.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mfx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mfx_op_vmult" in names and "mfx_cg_solve" in names and "mfx_bk5_apply_f64" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from benchmarks_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libmfx.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from benchmarks_amd import _lib
    assert sorted(_lib.exported_names()) == declared_functions()


def test_version_and_error_without_gpu():
    """Calls that touch no device: version and argument validation."""
    from benchmarks_amd import _lib
    L = _lib.lib()
    assert L.mfx_version() >= 100
    h = ctypes.c_void_p()
    assert L.mfx_op_create_box(None, ctypes.byref(h)) == 1  # MFX_ERR_INVALID_ARG
    assert b"null" in L.mfx_last_error()
