"""A reference-style C++ caller linked against libgstar4d.so through include/gdelaunay_shim.hpp
(the reference's gdelInit / gdelCompute / getTetraFromGpu / gdelDeInit) and run as its own
process: it must reorder the caller's points into Morton order like PointData::init
(GDelaunay/GDelaunay/GDelData.cu:125-130) and return the tetrahedra and adjacency of the C-ABI
call in that index space, whose canonical set equals the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import gstar4d_amd as g
from gstar4d_amd import datasets

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLER = r"""
#include <cstdio>
#include <vector>
#include "gdelaunay_shim.hpp"
using namespace gdel;

int main(int argc, char **argv)
{
    // argv: points.bin (N x 3 doubles), grid, out_points.bin, out_tets.bin
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    fseek(f, 0, SEEK_END);
    const long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    Point3HVec points((size_t)bytes / sizeof(Point3)), scaled;
    if (fread(points.data(), sizeof(Point3), points.size(), f) != points.size()) return 3;
    fclose(f);
    Config cfg;
    cfg._gridSize = atoi(argv[2]);
    try {
        gdelInit(cfg, points, scaled);
        double tPba, tStar, tSplay, tOut;
        gdelCompute(tPba, tStar, tSplay, tOut);
        TetraMesh mesh;
        getTetraFromGpu(mesh);
        gdelDeInit();
        FILE *o = fopen(argv[3], "wb");
        fwrite(points.data(), sizeof(Point3), points.size(), o);
        fclose(o);
        o = fopen(argv[4], "wb");
        fwrite(mesh._tetraVec.data(), sizeof(Tetrahedron), mesh._tetraVec.size(), o);
        fclose(o);
        printf("shim caller: %zu points, %zu tetrahedra, pba %.2f ms, star %.2f ms, splay %.2f ms, out %.2f ms\n",
               points.size(), mesh._tetraVec.size(), tPba, tStar, tSplay, tOut);
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
"""


def test_cpp_caller_through_shim(tmp_path, orc):
    grid = 64
    x = datasets.uniform_cube(20000, seed=21)
    src = tmp_path / "caller.cpp"
    src.write_text(CALLER)
    exe = tmp_path / "caller"
    lib_dir = os.path.join(ROOT, "gstar4d_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", lib_dir, "-lgstar4d", "-Wl,-rpath," + lib_dir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    pin = tmp_path / "points.bin"
    np.ascontiguousarray(x, "<f8").tofile(pin)
    pout, tout = tmp_path / "sorted.bin", tmp_path / "tets.bin"
    print("running the C++ caller on %d points, grid %d" % (x.shape[0], grid), flush=True)
    r = subprocess.run([str(exe), str(pin), str(grid), str(pout), str(tout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stderr
    sorted_pts = np.fromfile(pout, "<f8").reshape(-1, 3)
    rec = np.fromfile(tout, "<i4").reshape(-1, 8)
    tets, opp = rec[:, :4], rec[:, 4:]
    # the same call through the C ABI in Morton index space
    with g.Triangulator(grid_size=grid, index_space="morton") as t:
        ref_tets, ref_opp, perm = t.triangulate(x)
    assert np.array_equal(sorted_pts, x[perm])
    assert np.array_equal(tets, ref_tets) and np.array_equal(opp, ref_opp)
    oracle, _, _ = orc.triangulate(x, grid)
    assert (g.canonical(tets, perm) == oracle).all()
