"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU needed:
nothing here launches device work)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(sg_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.append(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "sg_uniform_cycle_dev" in names and "sg_last_error" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    from code_only_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_default_params_are_the_reference_values():
    from code_only_amd import _lib
    p = _lib.default_params()
    assert (p.tol, p.maxit, p.ncycle, p.npre, p.npost) == (1e-2, 10, 2, 2, 2)
    assert p.stop_mode == _lib.SG_STOP_REFERENCE
    assert _lib.load().sg_version() >= 1


def test_create_rejects_bad_sizes_without_touching_the_gpu():
    from code_only_amd import _lib
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    for n in (0, 3, 96, 4096):
        rc = lib.sg_uniform_create(ctypes.byref(ctx), n, 1.0, _lib.SG_PREC_FP64)
        assert rc == _lib.SG_ERR_ARG
        assert b"power of two" in lib.sg_last_error()
