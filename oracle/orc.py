"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The product library never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


class OrcOptions(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_int32), ("min_workset", ctypes.c_int32),
                ("max_rounds", ctypes.c_int32), ("verbose", ctypes.c_int32),
                ("insert_order", ctypes.c_int32)]


class OrcStats(ctypes.Structure):
    _fields_ = [("num_points", ctypes.c_int32), ("num_missing", ctypes.c_int32),
                ("num_grid_pairs", ctypes.c_int64), ("num_workset", ctypes.c_int64),
                ("max_workset", ctypes.c_int32), ("num_pumped", ctypes.c_int32),
                ("splay_rounds", ctypes.c_int32), ("splay_insertions", ctypes.c_int64),
                ("num_lowup_tetra", ctypes.c_int32), ("num_tetra", ctypes.c_int32),
                ("orient4_calls", ctypes.c_int64), ("exact_calls", ctypes.c_int64),
                ("sos_calls", ctypes.c_int64),
                ("t_morton", ctypes.c_double), ("t_pba", ctypes.c_double),
                ("t_star", ctypes.c_double), ("t_splay", ctypes.c_double),
                ("t_out", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint64)
_lp = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            from . import build as _b  # noqa: F401  (package-relative import when available)
        _LIB = ctypes.CDLL(path)
        L = _LIB
        L.orc_orient3d_exact.argtypes = [_dp] * 4
        L.orc_insphere_exact.argtypes = [_dp] * 5
        L.orc_orient3d.argtypes = [_dp] * 4
        L.orc_insphere.argtypes = [_dp] * 5
        L.orc_sign_minor.argtypes = [ctypes.c_int, ctypes.POINTER(_dp), _ip]
        L.orc_orient4_sos_xyz.argtypes = [_dp] * 5 + [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)]
        L.orc_orient4_sos.argtypes = [_dp] + [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)]
        L.orc_pred_counter.argtypes = [ctypes.c_int]
        L.orc_pred_counter.restype = ctypes.c_int64
        L.orc_morton_order.argtypes = [_dp, ctypes.c_int32, ctypes.c_int32, _ip, _ip]
        L.orc_scale_to_voxels.argtypes = [_dp, ctypes.c_int32, ctypes.c_int32, _ip]
        L.orc_pba.argtypes = [_ip, ctypes.c_int32, ctypes.c_int32, _ip]
        L.orc_pba_codes.argtypes = [_ip, ctypes.c_int32, ctypes.c_int32, _ip]
        L.orc_grid_pairs.argtypes = [_ip, ctypes.c_int32, _up, ctypes.c_int64]
        L.orc_grid_pairs.restype = ctypes.c_int64
        L.orc_triangulate.argtypes = [_dp, ctypes.c_int32, ctypes.POINTER(OrcOptions), _ip, _ip,
                                      ctypes.c_int64, ctypes.POINTER(OrcStats)]
        L.orc_triangulate.restype = ctypes.c_int64
        L.orc_triangulate_ex.argtypes = [_dp, ctypes.c_int32, ctypes.POINTER(OrcOptions), _up, ctypes.c_int64,
                                         _ip, _ip, ctypes.c_int64, ctypes.POINTER(OrcStats)]
        L.orc_triangulate_ex.restype = ctypes.c_int64
        L.orc_canonicalize.argtypes = [_ip, ctypes.c_int64, _ip]
        L.orc_adjacency.argtypes = [_ip, ctypes.c_int64, ctypes.c_int32, _ip]
        L.orc_validate.argtypes = [_dp, ctypes.c_int32, _ip, ctypes.c_int64, ctypes.c_int, _lp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_cube_pairs.argtypes = [_ip, _up]
    return _LIB


def ref_lib():
    """The reference's own predicates.c built in oracle/_ref (None if not built)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_predicates.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
        _REF.orient3d.argtypes = [_dp] * 4
        _REF.orient3d.restype = ctypes.c_double
        _REF.insphere.argtypes = [_dp] * 5
        _REF.insphere.restype = ctypes.c_double
        _REF.orient3dexact.argtypes = [_dp] * 4
        _REF.orient3dexact.restype = ctypes.c_double
        _REF.insphereexact.argtypes = [_dp] * 5
        _REF.insphereexact.restype = ctypes.c_double
        _REF.exactinit()
    return _REF


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _pt(p):
    a = np.ascontiguousarray(p, dtype=np.float64)
    return a, _p(a)


def orient3d(a, b, c, d, exact=False):
    arrs = [_pt(x) for x in (a, b, c, d)]
    f = lib().orc_orient3d_exact if exact else lib().orc_orient3d
    return f(*[x[1] for x in arrs])


def insphere(a, b, c, d, e, exact=False):
    arrs = [_pt(x) for x in (a, b, c, d, e)]
    f = lib().orc_insphere_exact if exact else lib().orc_insphere
    return f(*[x[1] for x in arrs])


def sign_minor(rows, cols):
    arrs = [_pt(r) for r in rows]
    rp = (_dp * len(rows))(*[x[1] for x in arrs])
    c = np.ascontiguousarray(cols, dtype=np.int32)
    return lib().orc_sign_minor(len(rows), rp, _p(c, _ip))


def orient4_sos(pts5, idx5):
    arrs = [_pt(x) for x in pts5]
    depth = ctypes.c_int(0)
    s = lib().orc_orient4_sos_xyz(*[x[1] for x in arrs], *[int(i) for i in idx5],
                                  ctypes.byref(depth))
    return s, depth.value


def morton_order(xyz, grid):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    perm = np.empty(n, np.int32)
    vox = np.empty((n, 3), np.int32)
    rc = lib().orc_morton_order(_p(xyz), n, grid, _p(perm, _ip), _p(vox, _ip))
    if rc != 0:
        raise ValueError("orc_morton_order failed")
    return perm, vox


def pba(vox_sorted, grid, codes=False):
    v = np.ascontiguousarray(vox_sorted, dtype=np.int32)
    out = np.empty(grid ** 3, np.int32)
    f = lib().orc_pba_codes if codes else lib().orc_pba
    if f(_p(v, _ip), v.shape[0], grid, _p(out, _ip)) != 0:
        raise ValueError("orc_pba failed")
    return out.reshape(grid, grid, grid)  # [z, y, x]


def cube_pairs(v8):
    """Pairs of one 2x2x2 cube of labels (corner c at offset (c&1, c>>1&1, c>>2)), in emission
    order, each as (min, max) with Marker = -1."""
    v = np.ascontiguousarray(v8, dtype=np.int32)
    out = np.zeros(32, np.uint64)
    m = lib().orc_cube_pairs(_p(v, _ip), _p(out, _up))
    return [(int(np.int32(np.uint32(k >> np.uint64(32)))), int(np.int32(np.uint32(k & np.uint64(0xFFFFFFFF)))))
            for k in out[:m]]


def grid_pairs(id_grid):
    g = id_grid.shape[0]
    grid = np.ascontiguousarray(id_grid, dtype=np.int32)
    m = lib().orc_grid_pairs(_p(grid, _ip), g, None, 0)
    out = np.empty(-m if m < 0 else max(m, 1), np.uint64)
    m = lib().orc_grid_pairs(_p(grid, _ip), g, _p(out, _up), out.shape[0])
    return out[:m]


def set_threads(nthreads):
    """Threads of the oracle's parallel stages (< 1: every processor); returns the count."""
    return lib().orc_set_threads(int(nthreads))


def triangulate(xyz, grid, min_workset=8, max_rounds=1000, canonical=True, threads=None, extra_pairs=None,
                insert_order=0):
    """Returns (tets, perm, stats).  tets are canonical (input index space, sorted) when
    canonical=True, else Morton-space with the reference's vertex order.  threads: OpenMP
    threads for this call (None keeps the current setting); the output does not depend on it.
    extra_pairs: uint64 (a << 32 | b) candidate pairs in Morton index space added to the
    worksets (the proposal's plug-in point).  insert_order: 0 inserts each star's list in
    ascending id (the reference's order), 1 nearest to the star point first."""
    if threads is not None:
        set_threads(threads)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    opt = OrcOptions(grid, min_workset, max_rounds, 0, insert_order)
    perm = np.empty(n, np.int32)
    st = OrcStats()
    cap = 8 * n + 64
    tets = np.empty((cap, 4), np.int32)
    ex = np.ascontiguousarray(extra_pairs if extra_pairs is not None else np.zeros(1, np.uint64), dtype=np.uint64)
    nex = 0 if extra_pairs is None else ex.shape[0]

    def call(t, c):
        return lib().orc_triangulate_ex(_p(xyz), n, ctypes.byref(opt), _p(ex, _up), nex, _p(perm, _ip),
                                        _p(t, _ip), c, ctypes.byref(st))
    cnt = call(tets, cap)
    if cnt < 0:
        raise ValueError("orc_triangulate failed with %d" % cnt)
    if cnt > cap:
        tets = np.empty((cnt, 4), np.int32)
        cnt = call(tets, cnt)
    tets = tets[:cnt]
    if canonical:
        lib().orc_canonicalize(_p(tets, _ip), cnt, _p(perm, _ip))
    return tets, perm, st.as_dict()


def canonicalize(tets, perm=None):
    t = np.ascontiguousarray(tets, dtype=np.int32).copy()
    lib().orc_canonicalize(_p(t, _ip), t.shape[0], None if perm is None else
                           _p(np.ascontiguousarray(perm, dtype=np.int32), _ip))
    return t


def adjacency(tets, n):
    t = np.ascontiguousarray(tets, dtype=np.int32)
    opp = np.empty_like(t)
    bad = lib().orc_adjacency(_p(t, _ip), t.shape[0], n, _p(opp, _ip))
    return opp, bad


def validate(xyz, tets, full_empty=True):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    t = np.ascontiguousarray(tets, dtype=np.int32)
    counts = np.zeros(4, np.int64)
    flags = lib().orc_validate(_p(xyz), xyz.shape[0], _p(t, _ip), t.shape[0], int(full_empty),
                               _p(counts, _lp))
    return flags, {"euler": int(counts[0]), "flat": int(counts[1]),
                   "adjacent_insphere": int(counts[2]), "empty_sphere": int(counts[3])}
