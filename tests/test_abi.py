"""CPU checks of the C-ABI boundary: the HIP library loads, exports every
entry point include/mocca_amd.h declares, and reports errors the documented
way.  No device compute is attempted (this container has no GPU)."""
import ctypes
import os
import re

import pytest

import mocca_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mocca_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mocca_[a-z0-9_]+)\s*\(", src)))


def test_header_and_bindings_agree():
    assert declared_symbols() == sorted(mocca_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = mocca_amd.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_invalid_arguments_fail_with_message():
    lib = mocca_amd.load()
    mesh = mocca_amd.MeshC(4, 4, 4, 0.8, 4.096, 1, 0)   # FD needs >= 6 points
    h = ctypes.c_void_p()
    rc = lib.mocca_solve_create(ctypes.byref(mesh), 8, 0, ctypes.byref(h))
    assert rc == -1
    assert b"at least 6 points" in lib.mocca_last_error(None)
    mesh = mocca_amd.MeshC(8, 8, 8, 0.8, 4.096, 1, 0)
    rc = lib.mocca_solve_create(ctypes.byref(mesh), 0, 0, ctypes.byref(h))
    assert rc == -1 and b"nwt" in lib.mocca_last_error(None)
    rc = lib.mocca_solve_create(ctypes.byref(mesh), 8, 7, ctypes.byref(h))
    assert rc == -1 and b"precision" in lib.mocca_last_error(None)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(mocca_amd, "LIBPATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(mocca_amd, "_LIB", None)
    with pytest.raises(mocca_amd.MoccaError):
        mocca_amd.load()
