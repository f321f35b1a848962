"""The C ABI library: it loads without a GPU and exports every function that
include/mgcm_rt.h and include/mgcm_sg.h declare; status strings; argument checks that need no device."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mgcm_rt.h")
HEADERS = (HEADER, os.path.join(ROOT, "include", "mgcm_sg.h"))


def declared_functions() -> list:
    """Every function include/mgcm_rt.h and include/mgcm_sg.h declare."""
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(mgcm_(?:rt|sg)_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("mgcm_rt_init", "mgcm_rt_destroy", "mgcm_rt_run", "mgcm_rt_run_host", "mgcm_rt_check",
              "mgcm_rt_run_batched", "mgcm_rt_planck_table", "mgcm_rt_laginterp", "mgcm_rt_strerror"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from mgcm_rt import load_library
    lib = load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_and_null_arguments():
    from mgcm_rt import load_library
    lib = load_library()
    assert lib.mgcm_rt_strerror(0) == b"ok" or lib.mgcm_rt_strerror(0)
    for st in (-1, -2, -3, -4, -5, -6):
        assert lib.mgcm_rt_strerror(st)
    # null arguments are rejected before any device call
    assert lib.mgcm_rt_init(None, None) == -1
    assert lib.mgcm_rt_run(None, None, None, 0, None) == -1
    assert lib.mgcm_rt_run_host(None, None, None, 0) == -1
    assert lib.mgcm_rt_destroy(None) in (0, -1)
    assert lib.mgcm_rt_set_tile(None, 1) == -1
    assert lib.mgcm_rt_tile_for(None, 10, 24) == -1


def test_python_binding_fails_loudly_without_library(tmp_path):
    import mgcm_rt
    saved = mgcm_rt._lib
    mgcm_rt._lib = None
    try:
        with pytest.raises(FileNotFoundError):
            mgcm_rt.load_library(str(tmp_path / "libmissing.so"))
    finally:
        mgcm_rt._lib = saved
