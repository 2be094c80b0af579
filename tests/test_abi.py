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


def test_python_call_sites_match_header_arity():
    """Every call("vx_...", ...) in the package, bench and entry point passes exactly the
    number of arguments include/voxhip.h declares (a starred MaskCache triple counts 3)."""
    import ast
    import glob
    decl = _lib.parse_header()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    files = (glob.glob(os.path.join(root, "voxurf_amd", "*.py")) +
             [os.path.join(root, "bench.py"), os.path.join(root, "__graft_entry__.py")])
    bad = []
    for f in files:
        tree = ast.parse(open(f).read())
        for node in ast.walk(tree):
            if not (isinstance(node, ast.Call) and getattr(node.func, "id", None) == "call"
                    and node.args and isinstance(node.args[0], ast.Constant)):
                continue
            name = node.args[0].value
            n = 0
            for a in node.args[1:]:
                n += 3 if isinstance(a, ast.Starred) else 1
            if name not in decl:
                bad.append((os.path.basename(f), node.lineno, name, "not declared"))
            elif n != len(decl[name][1]):
                bad.append((os.path.basename(f), node.lineno, name, n, len(decl[name][1])))
    assert not bad, bad
