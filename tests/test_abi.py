"""The C-ABI library loads here (no GPU) and exports every entry point include/gstar4d.h
declares; the C++ shim compiles.  No compute call is made without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gstar4d.h")
LIB = os.path.join(ROOT, "gstar4d_amd", "libgstar4d.so")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(g4d_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("g4d_create", "g4d_destroy", "g4d_triangulate", "g4d_triangulate_device",
                 "g4d_get_stats", "g4d_last_error", "g4d_copy_result"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libgstar4d.so is not built; run __graft_entry__.build()")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (g4d_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_python_binding_matches_exports():
    import gstar4d_amd as g
    L = g.load_library()
    for n in declared_functions():
        assert hasattr(L, n)
    assert set(g.EXPORTS) == set(declared_functions())


def test_default_options_without_device():
    import gstar4d_amd as g
    o = g.default_options()
    assert o.grid_size == 256 and o.min_workset == 8 and o.compute_adjacency == 1


def test_cpp_shim_compiles(tmp_path):
    src = tmp_path / "use_shim.cpp"
    src.write_text('#include "gdelaunay_shim.hpp"\n'
                   "int main(){ gdel::Config c; gdel::Point3HVec p, s; gdel::TetraMesh m;\n"
                   " (void)c; (void)p; (void)s; (void)m; return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_python_structs_match_c_sizes():
    import gstar4d_amd as g
    L = g.load_library()
    out = (ctypes.c_int64 * 6)()
    assert L.g4d_struct_sizes(out, 6) == 6
    mine = [ctypes.sizeof(c) for c in (g.Options, g.Result, g.Stats, g.CheckReport, g.PointFile, g.BatchResult)]
    assert list(out) == mine


def test_size_guard_limits():
    """g4d_check_sizes refuses lists the 32-bit sort offsets and face keys cannot hold."""
    import gstar4d_amd as g
    ok, cap = 0, 7
    assert g.check_sizes(0, 0) == ok
    assert g.check_sizes(2 ** 32 - 1, 0) == ok
    assert g.check_sizes(2 ** 32, 0) == cap
    assert g.check_sizes(0, 2 ** 30 - 1) == ok       # face key 4 t + 3 + 1 <= 2^32 - 1
    assert g.check_sizes(0, 2 ** 30) == cap
    assert g.check_sizes(-1, 0) == 1
    # uniform inputs: ~20 workset keys and ~6.8 tetrahedra per point
    n = 150_000_000
    assert g.check_sizes(20 * n, int(6.8 * n)) == ok
    n = 220_000_000
    assert g.check_sizes(20 * n, int(6.8 * n)) == cap


def test_batch_grids_length_is_checked():
    """A grids array shorter than the set count is refused before any device call."""
    import numpy as np
    import gstar4d_amd as g
    t = g.Triangulator.__new__(g.Triangulator)  # no device context: the argument check comes first
    t._L = g.load_library()
    t._ctx = None
    x = np.zeros((30, 3))
    with pytest.raises(ValueError):
        t.triangulate_batch(x, np.array([0, 10, 20, 30]), grids=np.array([16, 16]))
    with pytest.raises(ValueError):
        t.triangulate_batch_device(0, np.array([0, 10, 20, 30]), grids=np.array([16, 16, 16, 16]))
