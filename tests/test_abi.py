"""The C ABI: libvoxhip.so loads without a GPU and exports every symbol that
include/voxhip.h declares (no compute calls; runs on the CPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from voxurf_amd import _lib


def test_header_parses_and_all_symbols_exported():
    decls = _lib.parse_header()
    assert len(decls) > 10
    h = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in decls if not hasattr(h, n)]
    assert not missing, missing


def test_exports_match_header_exactly():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (vx_\w+)", out))
    assert exported == set(_lib.parse_header()), exported ^ set(_lib.parse_header())


def test_binding_sets_argtypes_and_error_path():
    lib = _lib.lib()
    assert lib.vx_version() >= 1
    # an argument error is reported through the status code and vx_last_error, without
    # touching the device
    rc = lib.vx_infer_t_minmax(0, None, None, None, None, 0.0, 1.0, -1, None, None, None)
    assert rc == 1
    assert b"n_rays" in lib.vx_last_error()
    rc = lib.vx_raw2alpha(7, None, 0.0, 1.0, 5, None, None, None)
    assert rc == 4 and b"dtype" in lib.vx_last_error()
