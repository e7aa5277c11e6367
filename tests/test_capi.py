"""The C-ABI library: it loads, exports every entry point include/pynof_amd.h
declares, and rejects bad arguments with error codes (no GPU needed)."""
import ctypes
import os
import re

import pytest

from pynof_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pynof_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pnjk_[a-z0-9_]+)\s*\(", txt)))


def test_library_built():
    assert _lib.lib_path().exists(), "run python -m pynof_amd.build"


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.declared_symbols())


def test_null_engine_errors():
    L = _lib.lib()
    rc = L.pnjk_jk_build(None, None, 0, 0, None, 0, 0, 0, None, None, None, None, None, None)
    assert rc == _lib.EINVAL
    assert b"engine" in L.pnjk_last_error()
    with pytest.raises(_lib.PnjkError):
        _lib.check(rc)
    rc = L.pnjk_jk_build_noalloc(None, None, 0, 0, None, 0, 0, 0, None, None, None, None, None, None)
    assert rc == _lib.EINVAL and b"engine" in L.pnjk_last_error()
    assert L.pnjk_upload_b(None, None, 1, 1, 0.0, None, None) == _lib.EINVAL
    assert L.pnjk_create(0, None) == _lib.EINVAL
    assert L.pnjk_destroy(None) == _lib.OK


def test_layout_helpers():
    L = _lib.lib()
    assert L.pnjk_packed_ld(24, 116) == 128
    assert L.pnjk_packed_ld(1000, 3600) == 3648
    ws = L.pnjk_workspace_bytes(1000, 3600, 300, 0)
    assert ws > 300 * 1000 * 3648 * 8  # the half-transformed tensor dominates
    assert L.pnjk_version().startswith(b"pynof_amd")
