"""The C-ABI library loads and exports every entry point its headers declare
(no GPU needed: nothing here launches a kernel)."""

import ctypes as C
import glob
import os
import re

import pytest

from vanguard_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(vg_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "vg_euler_solve_2d" in names and "vg_ctx_create" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_bindings_cover_the_headers():
    assert set(declared_functions()) <= set(_lib.SIGNATURES), "ctypes signatures missing for some entry points"


def test_invalid_arguments_are_rejected_without_a_device():
    lib = _lib.load()
    mp = _lib.EulerParams(1.4, 0, 25, 0, 0.25, 250)
    st = _lib.RunStats()
    rc = lib.vg_euler_solve_1d(None, None, None, 10, 0.1, 0.9, 0.1, 0, 0.0, 0, 2, 0, C.byref(mp), 0, None, 0,
                               C.byref(st))
    assert rc == _lib.VG_ERR_ARG
    assert b"context" in lib.vg_last_error()
    assert lib.vg_ctx_create(None, 0) == _lib.VG_ERR_ARG
    assert lib.vg_version().startswith(b"vanguard")
