"""The allocation-free entry point (g4d_triangulate_device_noalloc) through the C ABI: its
tetrahedra, adjacency and permutation equal those of g4d_triangulate_device + g4d_copy_result
on the same points, bit for bit, on two calls in a row (the second smaller, so stale scratch
contents from the first would show), without changing the device's free memory, and sizes
beyond its limits are refused with a clear error."""
import ctypes

import numpy as np
import pytest

import gstar4d_amd as g
from gstar4d_amd import datasets

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """The few HIP runtime calls the test needs as the caller: its own device buffers."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 16)) == 0
        return p.value

    def free(self, p):
        assert self.L.hipFree(ctypes.c_void_p(p)) == 0

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.malloc(arr.nbytes)
        assert self.L.hipMemcpy(ctypes.c_void_p(p), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, H2D) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        if out.nbytes:
            assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), out.nbytes, D2H) == 0
        return out

    def free_bytes(self):
        assert self.L.hipDeviceSynchronize() == 0
        f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
        assert self.L.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value


def _reference(grid, x):
    """The existing entry point: g4d_triangulate_device, then g4d_copy_result."""
    with g.Triangulator(grid_size=grid, adjacency=True) as t:
        d, n = t.upload(x)
        nt = t.triangulate_device(d, n)
        return t.copy_result(nt, n)


def test_noalloc_matches_triangulate_device_twice():
    hip = _Hip()
    grid = 64
    # the largest uniform set the call accepts, then a smaller degenerate-rich hard mix
    cases = [datasets.uniform_cube(g.NOALLOC_MAX_POINTS, seed=11), datasets.hard_mix(6000, seed=3)]
    refs = [_reference(grid, x) for x in cases]
    bufs = []
    for x in cases:
        n = x.shape[0]
        cap = 8 * n + 64
        bufs.append((hip.upload(x), hip.malloc(16 * cap), hip.malloc(16 * cap), hip.malloc(4 * n), cap))
    try:
        with g.Triangulator(grid_size=grid, adjacency=True) as t:
            # Warm-up calls: the HIP runtime allocates the private segment of a kernel that spills
            # registers at its first dispatch on a queue; after these, the device's free memory
            # measures allocations made by the call itself.
            for x, (d_xyz, d_tets, d_opp, d_perm, cap) in zip(cases, bufs):
                t.triangulate_device_noalloc(d_xyz, x.shape[0], d_tets, d_opp, cap, d_perm)
            for k, (x, (d_xyz, d_tets, d_opp, d_perm, cap)) in enumerate(zip(cases, bufs)):
                n = x.shape[0]
                print("noalloc case %d: %d points, grid %d" % (k, n, grid), flush=True)
                hip.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
                for p, nbytes in ((d_tets, 16 * cap), (d_opp, 16 * cap), (d_perm, 4 * n)):
                    assert hip.L.hipMemset(ctypes.c_void_p(p), 0xff, nbytes) == 0
                before = hip.free_bytes()
                nt, used = t.triangulate_device_noalloc(d_xyz, n, d_tets, d_opp, cap, d_perm)
                after = hip.free_bytes()
                print("noalloc case %d: %d tetrahedra, %d scratch bytes" % (k, nt, used), flush=True)
                assert after == before, "the allocation-free call changed the device's free memory"
                ref_tets, ref_opp, ref_perm = refs[k]
                assert nt == ref_tets.shape[0]
                assert np.array_equal(hip.download(d_tets, (nt, 4), np.int32), ref_tets)
                assert np.array_equal(hip.download(d_opp, (nt, 4), np.int32), ref_opp)
                assert np.array_equal(hip.download(d_perm, (n,), np.int32), ref_perm)
                assert 0 < used <= g.NOALLOC_SCRATCH_BYTES
    finally:
        for b in bufs:
            for p in b[:4]:
                hip.free(p)


def test_noalloc_limits_are_refused():
    hip = _Hip()
    x = datasets.uniform_cube(2000, seed=5)
    d_xyz = hip.upload(x)
    d_tets = hip.malloc(16 * 4)
    try:
        with g.Triangulator(grid_size=g.NOALLOC_MAX_GRID * 2, adjacency=False) as t:
            with pytest.raises(g.G4DError) as e:
                t.triangulate_device_noalloc(d_xyz, x.shape[0], d_tets, None, 4, None)
            assert e.value.status == "CAPACITY"
            t.set_grid(32)
            with pytest.raises(g.G4DError) as e:
                t.triangulate_device_noalloc(d_xyz, g.NOALLOC_MAX_POINTS + 1, d_tets, None, 4, None)
            assert e.value.status == "CAPACITY"
            # output too small: the count is reported, nothing written
            with pytest.raises(g.G4DError) as e:
                t.triangulate_device_noalloc(d_xyz, x.shape[0], d_tets, None, 4, None)
            assert e.value.status == "CAPACITY"
    finally:
        hip.free(d_xyz)
        hip.free(d_tets)
