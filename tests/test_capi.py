"""CPU-side checks of the C ABI: the library loads and exports exactly the
entry points include/gfmd_hip.h declares (no compute calls, no GPU)."""
import ctypes
import os
import subprocess

import pytest

import gfmd_hip


def test_library_built():
    assert os.path.exists(gfmd_hip.LIB_PATH), "run __graft_entry__.build()"


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(gfmd_hip.LIB_PATH)
    declared = gfmd_hip.header_functions()
    assert "gfmd_post_force" in declared and "gfmd_create" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing


def test_no_undeclared_gfmd_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", gfmd_hip.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines()
                if " T " in ln and ln.split()[-1].startswith("gfmd_")}
    assert exported == set(gfmd_hip.header_functions())


def test_version_and_error_string_without_gpu():
    L = gfmd_hip.load()
    assert b"gfx950" in L.gfmd_version()
    assert L.gfmd_last_error() == b""


def test_argument_errors_are_status_codes():
    """Invalid arguments are rejected before any device call."""
    L = gfmd_hip.load()
    h = ctypes.c_void_p()
    assert L.gfmd_create(0, 8, 3, 0, ctypes.byref(h)) == gfmd_hip.ERR_ARG
    assert b"invalid grid" in L.gfmd_last_error()
    assert L.gfmd_create(8, 8, 4, 0, ctypes.byref(h)) == gfmd_hip.ERR_ARG
    assert L.gfmd_create(8, 8, 3, 7, ctypes.byref(h)) == gfmd_hip.ERR_ARG
    assert L.gfmd_destroy(None) == gfmd_hip.OK
    with pytest.raises(gfmd_hip.GFMDError):
        gfmd_hip.check(L.gfmd_create(8, 8, 30, 0, ctypes.byref(h)))
