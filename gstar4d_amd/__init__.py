"""gStar4D for MI355X — host-side Python interface over the C ABI (include/gstar4d.h).

The reference's triangulation interface (GDelaunay/GDelaunay/GDelaunay.h:39-42) is mirrored
by gdel_init / gdel_compute / get_tetra / gdel_deinit; Triangulator is the object form.
All computation runs in libgstar4d.so (hand-written gfx950 HIP kernels); when the library
is missing this module raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

__all__ = ["Triangulator", "Options", "Stats", "CheckReport", "G4DError", "load_library", "gdel_init",
           "gdel_compute", "get_tetra", "gdel_deinit", "canonical", "LIB_PATH", "PointFile", "read_asc",
           "write_asc", "write_ply", "BatchResult", "batch_grid_size", "check_sizes"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgstar4d.so")

# Limits of the allocation-free entry point (include/gstar4d.h).
NOALLOC_SCRATCH_BYTES = 256 << 20
NOALLOC_MAX_POINTS = 16384
NOALLOC_MAX_GRID = 128

STATUS = {0: "OK", 1: "INVALID_INPUT", 2: "DUPLICATE_POINTS", 3: "GRID_LIMIT", 4: "OOM",
          5: "DEVICE", 6: "NOT_CONVERGED", 7: "CAPACITY", 8: "PREDICATE_RANGE", 9: "INTERNAL"}


class G4DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code
        self.status = STATUS.get(code, "?")


class Options(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_int32), ("min_workset", ctypes.c_int32),
                ("max_rounds", ctypes.c_int32), ("index_space", ctypes.c_int32),
                ("device", ctypes.c_int32), ("compute_adjacency", ctypes.c_int32),
                ("lds_star_limit", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("num_tets", ctypes.c_int64), ("tets", ctypes.POINTER(ctypes.c_int32)),
                ("opp", ctypes.POINTER(ctypes.c_int32)), ("perm", ctypes.POINTER(ctypes.c_int32)),
                ("num_points", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("t_morton_ms", ctypes.c_double), ("t_pba_ms", ctypes.c_double),
                ("t_pairs_ms", ctypes.c_double), ("t_star_ms", ctypes.c_double),
                ("t_splay_ms", ctypes.c_double), ("t_out_ms", ctypes.c_double),
                ("t_adj_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double),
                ("num_points", ctypes.c_int32), ("grid_size", ctypes.c_int32),
                ("splay_rounds", ctypes.c_int32), ("max_workset", ctypes.c_int32),
                ("big_stars", ctypes.c_int32), ("huge_stars", ctypes.c_int32),
                ("num_missing", ctypes.c_int64), ("num_grid_pairs", ctypes.c_int64),
                ("num_workset", ctypes.c_int64), ("num_pumped", ctypes.c_int64),
                ("splay_insertions", ctypes.c_int64), ("num_tets", ctypes.c_int64),
                ("exact_calls", ctypes.c_int64), ("sos_calls", ctypes.c_int64),
                ("exact_stars", ctypes.c_int64),
                ("pba_bytes", ctypes.c_int64), ("t_pba_pass_ms", ctypes.c_double * 3),
                ("star_store_bytes", ctypes.c_int64), ("store_repacks", ctypes.c_int32),
                ("star_passes", ctypes.c_int32), ("star_t0_launches", ctypes.c_int32), ("pad1", ctypes.c_int32),
                ("t_star_t0_ms", ctypes.c_double), ("star_t0_filter_evals", ctypes.c_int64),
                ("star_t0_cofactors", ctypes.c_int64), ("prop_enabled", ctypes.c_int32), ("pba_bands", ctypes.c_int32),
                ("prop_pairs", ctypes.c_int64), ("t_prop_ms", ctypes.c_double),
                ("graph_launches", ctypes.c_int32), ("graph_captures", ctypes.c_int32)]

    def as_dict(self):
        d = {}
        for f, _ in self._fields_:
            v = getattr(self, f)
            d[f] = list(v) if f == "t_pba_pass_ms" else v
        return d


class PointFile(ctypes.Structure):
    """g4d_point_file: points of an .asc file (GD/Main/Application.cpp:475-585)."""
    _fields_ = [("num_points", ctypes.c_int32), ("pad", ctypes.c_int32), ("num_read", ctypes.c_int64),
                ("num_duplicates", ctypes.c_int64), ("min_val", ctypes.c_float), ("max_val", ctypes.c_float),
                ("points", ctypes.POINTER(ctypes.c_double)), ("scaled", ctypes.POINTER(ctypes.c_double))]


class BatchResult(ctypes.Structure):
    """g4d_batch_result."""
    _fields_ = [("num_sets", ctypes.c_int32), ("pad", ctypes.c_int32), ("num_points", ctypes.c_int64),
                ("num_tets", ctypes.c_int64), ("tet_offsets", ctypes.POINTER(ctypes.c_int64)),
                ("tets", ctypes.POINTER(ctypes.c_int32)), ("opp", ctypes.POINTER(ctypes.c_int32)),
                ("grid_sizes", ctypes.POINTER(ctypes.c_int32))]


class CheckReport(ctypes.Structure):
    """g4d_check_report: TetraMesh::check (GD/Common/Geometry.cu:200-393) plus adjacency and
    empty-sphere counts."""
    _fields_ = [(f, ctypes.c_int64) for f in (
        "num_tets", "num_vertices", "num_edges", "num_faces", "hull_faces", "euler", "orientation_failures",
        "flat_tets", "insphere_failures", "adjacency_failures", "empty_sphere_failures", "empty_sphere_tests")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_LIB = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint64)
_lp = ctypes.POINTER(ctypes.c_int64)

EXPORTS = ["g4d_default_options", "g4d_create", "g4d_destroy", "g4d_last_error", "g4d_device_count",
           "g4d_triangulate", "g4d_triangulate_device", "g4d_copy_result", "g4d_upload_points",
           "g4d_get_stats", "g4d_set_options", "g4d_diag_orient4", "g4d_diag_orient3d",
           "g4d_diag_morton", "g4d_diag_pba", "g4d_diag_grid_pairs", "g4d_diag_copy",
           "g4d_check_result", "g4d_check_mesh", "g4d_read_asc", "g4d_free_point_file", "g4d_write_asc",
           "g4d_write_ply", "g4d_triangulate_batch", "g4d_triangulate_batch_device", "g4d_batch_grid_size",
           "g4d_struct_sizes", "g4d_check_sizes", "g4d_set_proposal", "g4d_set_sample_export", "g4d_get_samples",
           "g4d_mlp_forward", "g4d_mlp_grad", "g4d_mlp_train", "g4d_check_stars", "g4d_diag_set_link_vertex",
           "g4d_triangulate_device_noalloc", "g4d_set_star_kernel", "g4d_set_grid_table", "g4d_grid_for_set",
           "g4d_set_pba_bands", "g4d_set_stage_graphs", "g4d_diag_sort_pairs"]

MLP_PARAMS = 7312
PROP_FEATURES = 12
PROP_SLOTS = 8
PROP_BINS = 8


def load_library(path=None):
    """Loads libgstar4d.so (building nothing); raises if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError("libgstar4d.so not built (run __graft_entry__.build()); no fallback exists")
    L = ctypes.CDLL(path)
    L.g4d_default_options.argtypes = [ctypes.POINTER(Options)]
    L.g4d_create.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(Options)]
    L.g4d_destroy.argtypes = [_vp]
    L.g4d_destroy.restype = None
    L.g4d_last_error.argtypes = [_vp]
    L.g4d_last_error.restype = ctypes.c_char_p
    L.g4d_triangulate.argtypes = [_vp, _dp, ctypes.c_int32, ctypes.POINTER(Result)]
    L.g4d_triangulate_device.argtypes = [_vp, _vp, ctypes.c_int32, _lp]
    L.g4d_copy_result.argtypes = [_vp, _ip, _ip, _ip]
    L.g4d_triangulate_device_noalloc.argtypes = [_vp, _vp, ctypes.c_int32, _vp, _vp, ctypes.c_int64, _vp, _lp, _lp]
    L.g4d_upload_points.argtypes = [_vp, _dp, ctypes.c_int32, ctypes.POINTER(_vp)]
    L.g4d_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
    L.g4d_set_options.argtypes = [_vp, ctypes.POINTER(Options)]
    L.g4d_diag_orient4.argtypes = [_vp, _dp, _ip, ctypes.c_int64, ctypes.c_int32, _ip, _ip]
    L.g4d_diag_orient3d.argtypes = [_vp, _dp, ctypes.c_int64, _ip]
    L.g4d_diag_morton.argtypes = [_vp, _dp, ctypes.c_int32, ctypes.c_int32, _ip, _ip]
    L.g4d_diag_pba.argtypes = [_vp, _ip, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _ip]
    L.g4d_diag_grid_pairs.argtypes = [_vp, _ip, ctypes.c_int32, _up, ctypes.c_int64, _lp]
    L.g4d_diag_copy.argtypes = [_vp, ctypes.c_int64, _dp]
    L.g4d_diag_sort_pairs.argtypes = [_vp, _up, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _up, _dp, _ip]
    L.g4d_check_result.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(CheckReport)]
    L.g4d_check_stars.argtypes = [_vp, _lp, _lp]
    L.g4d_diag_set_link_vertex.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.g4d_check_mesh.argtypes = [_vp, _dp, ctypes.c_int32, _ip, _ip, ctypes.c_int64, ctypes.c_int32,
                                 ctypes.POINTER(CheckReport)]
    L.g4d_triangulate_batch.argtypes = [_vp, _dp, _lp, ctypes.c_int32, _ip, ctypes.POINTER(BatchResult)]
    L.g4d_triangulate_batch_device.argtypes = [_vp, _vp, _lp, ctypes.c_int32, _ip, _lp]
    L.g4d_batch_grid_size.argtypes = [ctypes.c_int64]
    L.g4d_set_star_kernel.argtypes = [_vp, ctypes.c_int32]
    L.g4d_set_grid_table.argtypes = [_vp, ctypes.c_int32, _lp, ctypes.POINTER(ctypes.c_float)]
    L.g4d_grid_for_set.argtypes = [_vp, ctypes.c_int64]
    L.g4d_set_pba_bands.argtypes = [_vp, ctypes.c_int32]
    L.g4d_set_stage_graphs.argtypes = [_vp, ctypes.c_int32]
    L.g4d_struct_sizes.argtypes = [_lp, ctypes.c_int32]
    L.g4d_check_sizes.argtypes = [ctypes.c_int64, ctypes.c_int64]
    _fp = ctypes.POINTER(ctypes.c_float)
    _u32p = ctypes.POINTER(ctypes.c_uint32)
    L.g4d_set_proposal.argtypes = [_vp, _fp, ctypes.c_int32, ctypes.c_float, ctypes.c_int32]
    L.g4d_set_sample_export.argtypes = [_vp, ctypes.c_int32]
    L.g4d_get_samples.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, _fp, _u32p, _ip, ctypes.c_int64, _lp]
    L.g4d_mlp_forward.argtypes = [_vp, _fp, _fp, ctypes.c_int64, _fp]
    L.g4d_mlp_grad.argtypes = [_vp, _fp, _fp, _u32p, ctypes.c_int64, _fp, ctypes.POINTER(ctypes.c_double)]
    L.g4d_mlp_train.argtypes = [_vp, _fp, _fp, _u32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.g4d_read_asc.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(PointFile)]
    L.g4d_free_point_file.argtypes = [ctypes.POINTER(PointFile)]
    L.g4d_free_point_file.restype = None
    L.g4d_write_asc.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int32]
    L.g4d_write_ply.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int32, _ip, ctypes.c_int64]
    _LIB = L
    return L


def read_asc(path, grid_size):
    """Reads an .asc point file as the reference's Application::readPoints does.  Returns
    (points, scaled, info): (N,3) float64 arrays and a dict with num_read, num_duplicates,
    min_val, max_val."""
    L = load_library()
    f = PointFile()
    rc = L.g4d_read_asc(os.fsencode(path), int(grid_size), ctypes.byref(f))
    try:
        if rc != 0:
            raise G4DError(rc, "g4d_read_asc(%s) failed" % path)
        n = f.num_points
        pts = np.ctypeslib.as_array(f.points, shape=(max(n, 1) * 3,))[:3 * n].reshape(n, 3).copy()
        sc = np.ctypeslib.as_array(f.scaled, shape=(max(n, 1) * 3,))[:3 * n].reshape(n, 3).copy()
        info = {"num_read": f.num_read, "num_duplicates": f.num_duplicates, "min_val": f.min_val,
                "max_val": f.max_val}
    finally:
        L.g4d_free_point_file(ctypes.byref(f))
    return pts, sc, info


def check_sizes(sort_keys, num_tets):
    """Status of g4d_check_sizes: 0 within the library's list and result limits, 7 (CAPACITY)
    beyond them."""
    return int(load_library().g4d_check_sizes(int(sort_keys), int(num_tets)))


def _grids_arg(grids, nsets):
    if grids is None:
        return None
    gr = np.ascontiguousarray(grids, dtype=np.int32)
    if gr.ndim != 1 or gr.shape[0] != nsets:
        raise ValueError("grids must hold one grid width per set (%d), got shape %s" % (nsets, gr.shape))
    return gr


def batch_grid_size(n):
    """The PBA grid width the batched mode picks for a set of n points."""
    return int(load_library().g4d_batch_grid_size(int(n)))


def write_asc(path, xyz):
    L = load_library()
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    rc = L.g4d_write_asc(os.fsencode(path), _ptr(x, _dp), x.shape[0])
    if rc != 0:
        raise G4DError(rc, "g4d_write_asc(%s) failed" % path)


def write_ply(path, xyz, tets):
    L = load_library()
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    rc = L.g4d_write_ply(os.fsencode(path), _ptr(x, _dp), x.shape[0], _ptr(t, _ip), t.shape[0])
    if rc != 0:
        raise G4DError(rc, "g4d_write_ply(%s) failed" % path)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def default_options(**kw):
    L = load_library()
    o = Options()
    L.g4d_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o


class Triangulator:
    """One device context.  triangulate(xyz) returns (tets, opp, perm): tets (T,4) int32 with
    orient3d(v0,v1,v2,v3) < 0, opp (T,4) neighbour across the face opposite each vertex or -1,
    perm (N,) Morton rank -> input index."""

    def __init__(self, grid_size=256, min_workset=8, max_rounds=1000, index_space="input",
                 device=0, adjacency=True, lds_star_limit=0):
        self._L = load_library()
        self.opt = default_options(grid_size=grid_size, min_workset=min_workset,
                                   max_rounds=max_rounds,
                                   index_space=0 if index_space == "input" else 1,
                                   device=device, compute_adjacency=1 if adjacency else 0,
                                   lds_star_limit=lds_star_limit)
        self._ctx = _vp()
        rc = self._L.g4d_create(ctypes.byref(self._ctx), ctypes.byref(self.opt))
        if rc != 0:
            raise G4DError(rc, "g4d_create failed (no HIP device?)")

    def close(self):
        if self._ctx:
            self._L.g4d_destroy(self._ctx)
            self._ctx = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise G4DError(rc, self._L.g4d_last_error(self._ctx).decode())

    def set_grid(self, grid_size):
        self.opt.grid_size = int(grid_size)
        self._check(self._L.g4d_set_options(self._ctx, ctypes.byref(self.opt)))

    def triangulate(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("xyz must be (N, 3)")
        res = Result()
        self._check(self._L.g4d_triangulate(self._ctx, _ptr(xyz, _dp), xyz.shape[0], ctypes.byref(res)))
        nt = res.num_tets
        tets = np.ctypeslib.as_array(res.tets, shape=(max(nt, 1), 4))[:nt].copy()
        opp = (np.ctypeslib.as_array(res.opp, shape=(max(nt, 1), 4))[:nt].copy()
               if res.opp else None)
        perm = np.ctypeslib.as_array(res.perm, shape=(xyz.shape[0],)).copy()
        return tets, opp, perm

    def triangulate_batch(self, xyz, set_offsets, grids=None):
        """Independent point sets in one pass of each stage.  set_offsets (S+1,) int64 with
        set_offsets[0] = 0.  Returns (tets, opp, tet_offsets, grid_sizes): set k's tetrahedra are
        tets[tet_offsets[k]:tet_offsets[k+1]] with vertex ids into that set's points and opp ids
        within that range (-1 on the hull)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        off = np.ascontiguousarray(set_offsets, dtype=np.int64)
        S = off.shape[0] - 1
        gr = _grids_arg(grids, S)
        res = BatchResult()
        self._check(self._L.g4d_triangulate_batch(self._ctx, _ptr(xyz, _dp), _ptr(off, _lp), S,
                                                  None if gr is None else _ptr(gr, _ip), ctypes.byref(res)))
        nt = res.num_tets
        tets = np.ctypeslib.as_array(res.tets, shape=(max(nt, 1), 4))[:nt].copy()
        opp = (np.ctypeslib.as_array(res.opp, shape=(max(nt, 1), 4))[:nt].copy() if res.opp else None)
        toff = np.ctypeslib.as_array(res.tet_offsets, shape=(S + 1,)).copy()
        gs = np.ctypeslib.as_array(res.grid_sizes, shape=(S,)).copy()
        return tets, opp, toff, gs

    def triangulate_batch_device(self, d_xyz, set_offsets, grids=None):
        off = np.ascontiguousarray(set_offsets, dtype=np.int64)
        gr = _grids_arg(grids, off.shape[0] - 1)
        nt = ctypes.c_int64(0)
        self._check(self._L.g4d_triangulate_batch_device(self._ctx, _vp(d_xyz), _ptr(off, _lp), off.shape[0] - 1,
                                                         None if gr is None else _ptr(gr, _ip), ctypes.byref(nt)))
        return nt.value

    def upload(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        d = _vp()
        self._check(self._L.g4d_upload_points(self._ctx, _ptr(xyz, _dp), xyz.shape[0], ctypes.byref(d)))
        return d.value, xyz.shape[0]

    def triangulate_device(self, d_xyz, n):
        nt = ctypes.c_int64(0)
        self._check(self._L.g4d_triangulate_device(self._ctx, _vp(d_xyz), int(n), ctypes.byref(nt)))
        return nt.value

    def set_star_kernel(self, variant=-1):
        """Variant of the 32-vertex star kernel (-1: built-in default; see g4d_set_star_kernel)."""
        self._check(self._L.g4d_set_star_kernel(self._ctx, int(variant)))

    def set_grid_table(self, max_points=(), voxels_per_point=()):
        """Grids of batched sets without explicit grids from a per-size table (empty: the
        built-in formula)."""
        mp = np.ascontiguousarray(max_points, dtype=np.int64)
        vp = np.ascontiguousarray(voxels_per_point, dtype=np.float32)
        if mp.shape != vp.shape:
            raise ValueError("max_points and voxels_per_point differ in length")
        self._check(self._L.g4d_set_grid_table(self._ctx, int(mp.shape[0]), _ptr(mp, _lp) if mp.size else None,
                                               _ptr(vp, ctypes.POINTER(ctypes.c_float)) if vp.size else None))

    def set_pba_bands(self, bands=0):
        """Bands of the y and x PBA passes (0: by grid width; 1: whole columns)."""
        self._check(self._L.g4d_set_pba_bands(self._ctx, int(bands)))

    def set_stage_graphs(self, mode=-1):
        """Grid stage as a replayed hipGraph: 1 always, 0 never, -1 batched calls only."""
        self._check(self._L.g4d_set_stage_graphs(self._ctx, int(mode)))

    def grid_for_set(self, n):
        return int(self._L.g4d_grid_for_set(self._ctx, int(n)))

    def triangulate_device_noalloc(self, d_xyz, n, d_tets, d_opp, max_tets, d_perm=None):
        """Allocation-free triangulation of n device-resident points into the caller's device
        buffers d_tets / d_opp (max_tets x 4 int32 each; d_opp may be None without adjacency)
        and d_perm (n int32, optional).  Returns (num_tets, scratch bytes used)."""
        nt, used = ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(self._L.g4d_triangulate_device_noalloc(self._ctx, _vp(d_xyz), int(n), _vp(d_tets),
                                                           _vp(d_opp) if d_opp else None, int(max_tets),
                                                           _vp(d_perm) if d_perm else None, ctypes.byref(nt),
                                                           ctypes.byref(used)))
        return nt.value, used.value

    def copy_result(self, nt, n, want_opp=True):
        tets = np.empty((max(nt, 1), 4), np.int32)
        opp = np.empty((max(nt, 1), 4), np.int32) if want_opp else None
        perm = np.empty(n, np.int32)
        self._check(self._L.g4d_copy_result(self._ctx, _ptr(tets, _ip),
                                            _ptr(opp, _ip) if want_opp else None, _ptr(perm, _ip)))
        return tets[:nt], (opp[:nt] if want_opp else None), perm

    def check(self, full_empty_sphere=False):
        """Validate the last result on the GPU; returns the report as a dict."""
        rep = CheckReport()
        self._check(self._L.g4d_check_result(self._ctx, int(full_empty_sphere), ctypes.byref(rep)))
        return rep.as_dict()

    def check_stars(self):
        """Star consistency of the last result's stars on the GPU: (inconsistent link
        triangles, link triangles checked)."""
        bad = ctypes.c_int64(0)
        tot = ctypes.c_int64(0)
        self._check(self._L.g4d_check_stars(self._ctx, ctypes.byref(bad), ctypes.byref(tot)))
        return int(bad.value), int(tot.value)

    def diag_set_link_vertex(self, star, local, v):
        """Overwrites one link vertex of a star of the last result (planted defects)."""
        self._check(self._L.g4d_diag_set_link_vertex(self._ctx, int(star), int(local), int(v)))

    def check_mesh(self, xyz, tets, opp=None, full_empty_sphere=True):
        """Validate any tetrahedral mesh of the points xyz on the GPU."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        t = np.ascontiguousarray(tets, dtype=np.int32)
        o = None if opp is None else np.ascontiguousarray(opp, dtype=np.int32)
        rep = CheckReport()
        self._check(self._L.g4d_check_mesh(self._ctx, _ptr(xyz, _dp), xyz.shape[0], _ptr(t, _ip),
                                           None if o is None else _ptr(o, _ip), t.shape[0],
                                           int(full_empty_sphere), ctypes.byref(rep)))
        return rep.as_dict()

    # ---- learned proposal ----
    def set_proposal(self, params=None, threshold=0.0, enable=True):
        """Turns the learned proposal on with `params` (G4D_MLP_PARAMS floats) or off."""
        if not enable:
            self._check(self._L.g4d_set_proposal(self._ctx, None, 0, 0.0, 0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        self._check(self._L.g4d_set_proposal(self._ctx, _ptr(p, ctypes.POINTER(ctypes.c_float)), p.shape[0],
                                              float(threshold), 1))

    def set_sample_export(self, enable=True):
        self._check(self._L.g4d_set_sample_export(self._ctx, int(bool(enable))))

    def get_samples(self, stride=1, offset=0, candidates=False):
        """(features (k,12) float32, labels (k,) uint32[, candidates (k,8) int32]) of the last call."""
        n = self.stats()["num_points"]
        cap = max(0, (n - offset + stride - 1) // stride)
        f = np.empty((max(cap, 1), PROP_FEATURES), np.float32)
        y = np.empty(max(cap, 1), np.uint32)
        c = np.empty((max(cap, 1), PROP_SLOTS), np.int32) if candidates else None
        cnt = ctypes.c_int64(0)
        self._check(self._L.g4d_get_samples(self._ctx, int(offset), int(stride),
                                             _ptr(f, ctypes.POINTER(ctypes.c_float)),
                                             _ptr(y, ctypes.POINTER(ctypes.c_uint32)),
                                             _ptr(c, _ip) if candidates else None, cap, ctypes.byref(cnt)))
        k = cnt.value
        return (f[:k], y[:k], c[:k]) if candidates else (f[:k], y[:k])

    def mlp_forward(self, params, features):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PROP_FEATURES)
        out = np.empty((x.shape[0], 16), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.g4d_mlp_forward(self._ctx, _ptr(p, fp), _ptr(x, fp), x.shape[0], _ptr(out, fp)))
        return out

    def mlp_grad(self, params, features, labels):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PROP_FEATURES)
        y = np.ascontiguousarray(labels, dtype=np.uint32).reshape(-1)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_double(0.0)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.g4d_mlp_grad(self._ctx, _ptr(p, fp), _ptr(x, fp), _ptr(y, ctypes.POINTER(ctypes.c_uint32)),
                                         x.shape[0], _ptr(g, fp), ctypes.byref(loss)))
        return g, loss.value

    def mlp_train(self, params, features, labels, epochs=4, batch=4096, lr=2e-3, seed=0):
        """Trains a copy of params on the GPU; returns (params, mean loss per epoch)."""
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1).copy()
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PROP_FEATURES)
        y = np.ascontiguousarray(labels, dtype=np.uint32).reshape(-1)
        el = np.zeros(epochs, np.float64)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.g4d_mlp_train(self._ctx, _ptr(p, fp), _ptr(x, fp), _ptr(y, ctypes.POINTER(ctypes.c_uint32)),
                                          x.shape[0], int(epochs), int(batch), float(lr), int(seed),
                                          _ptr(el, ctypes.POINTER(ctypes.c_double))))
        return p, el

    def stats(self):
        st = Stats()
        self._check(self._L.g4d_get_stats(self._ctx, ctypes.byref(st)))
        return st.as_dict()

    # ---- diagnostics ----
    def diag_orient4(self, pts5, idx5, mode=0):
        pts5 = np.ascontiguousarray(pts5, dtype=np.float64).reshape(-1, 15)
        idx5 = np.ascontiguousarray(idx5, dtype=np.int32).reshape(-1, 5)
        T = pts5.shape[0]
        sign = np.zeros(T, np.int32)
        depth = np.zeros(T, np.int32)
        self._check(self._L.g4d_diag_orient4(self._ctx, _ptr(pts5, _dp), _ptr(idx5, _ip), T, mode,
                                             _ptr(sign, _ip), _ptr(depth, _ip)))
        return sign, depth

    def diag_orient3d(self, pts4):
        pts4 = np.ascontiguousarray(pts4, dtype=np.float64).reshape(-1, 12)
        sign = np.zeros(pts4.shape[0], np.int32)
        self._check(self._L.g4d_diag_orient3d(self._ctx, _ptr(pts4, _dp), pts4.shape[0], _ptr(sign, _ip)))
        return sign

    def diag_morton(self, xyz, grid):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = xyz.shape[0]
        perm = np.empty(n, np.int32)
        vox = np.empty((n, 3), np.int32)
        self._check(self._L.g4d_diag_morton(self._ctx, _ptr(xyz, _dp), n, grid, _ptr(perm, _ip), _ptr(vox, _ip)))
        return perm, vox

    def diag_pba(self, vox_sorted, grid, codes=False, id_grids=False):
        """Nearest-site grid [z, y, x]: point ids, or site codes with codes=True; id_grids=True
        runs the point-id grid kernels (used above 1024) at any grid width."""
        v = np.ascontiguousarray(vox_sorted, dtype=np.int32)
        out = np.empty(grid ** 3, np.int32)
        mode = 2 if id_grids else int(bool(codes))
        self._check(self._L.g4d_diag_pba(self._ctx, _ptr(v, _ip), v.shape[0], grid, mode, _ptr(out, _ip)))
        return out.reshape(grid, grid, grid)

    def diag_grid_pairs(self, id_grid):
        g = id_grid.shape[0]
        grid = np.ascontiguousarray(id_grid, dtype=np.int32)
        cnt = ctypes.c_int64(0)
        self._check(self._L.g4d_diag_grid_pairs(self._ctx, _ptr(grid, _ip), g, None, 0, ctypes.byref(cnt)))
        out = np.empty(max(cnt.value, 1), np.uint64)
        self._check(self._L.g4d_diag_grid_pairs(self._ctx, _ptr(grid, _ip), g, _ptr(out, _up), out.shape[0],
                                                ctypes.byref(cnt)))
        return out[:cnt.value]

    def diag_sort_pairs(self, keys, nb, method=0):
        """Sorts pair keys a << 32 | b (a, b < 2^nb) on the device: method 0 the pipeline's
        choice, 1 the LSD radix sort, 3 the bucket sort at any length; 2 and 4 the same as 0 and
        3 for keys a << 33 | b with b < 2^(nb+1).  Returns (sorted keys, ms, bucket shift)."""
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty(max(k.shape[0], 1), np.uint64)
        ms, sh = ctypes.c_double(0.0), ctypes.c_int32(0)
        self._check(self._L.g4d_diag_sort_pairs(self._ctx, _ptr(k, _up), k.shape[0], int(nb), int(method),
                                                _ptr(out, _up), ctypes.byref(ms), ctypes.byref(sh)))
        return out[:k.shape[0]], ms.value, sh.value

    def diag_copy(self, nbytes):
        """Plain device copy of nbytes; returns the kernel time in ms."""
        ms = ctypes.c_double(0.0)
        self._check(self._L.g4d_diag_copy(self._ctx, int(nbytes), ctypes.byref(ms)))
        return ms.value


def canonical(tets, perm=None):
    """Canonical tetra set: optional index map, each quadruple sorted, rows sorted."""
    t = np.asarray(tets, dtype=np.int64)
    if perm is not None:
        t = np.asarray(perm, dtype=np.int64)[t]
    t = np.sort(t, axis=1)
    order = np.lexsort((t[:, 3], t[:, 2], t[:, 1], t[:, 0]))
    return t[order].astype(np.int32)


# ---- the reference's four-call interface (module-level state, like GDelHost.cu's globals) ----
_STATE = {"tri": None, "points": None, "result": None, "times": None}


def gdel_init(config, points, scaled=None):
    """gdelInit (GD/GDelaunay/GDelaunay.cpp:47-60).  config: dict with grid_size (and optional
    min_workset / device).  Like the reference, `points` (an (N,3) float64 array) is reordered
    in place into Morton order when config['sort_in_place'] is true (the default)."""
    gdel_deinit()
    tri = Triangulator(grid_size=config.get("grid_size", 256), min_workset=config.get("min_workset", 8),
                       index_space="morton", device=config.get("device", 0))
    _STATE.update(tri=tri, points=points, result=None, times=None,
                  sort=config.get("sort_in_place", True))


def gdel_compute():
    """gdelCompute: returns (t_pba, t_star, t_splay, t_out) in ms."""
    tri = _STATE["tri"]
    if tri is None:
        raise RuntimeError("gdel_init not called")
    pts = _STATE["points"]
    tets, opp, perm = tri.triangulate(pts)
    if _STATE.get("sort", True):
        pts[:] = pts[perm]  # GD/GDelaunay/GDelData.cu:125-130 rewrites the caller's points
    _STATE["result"] = (tets, opp, perm)
    st = tri.stats()
    _STATE["times"] = (st["t_pba_ms"], st["t_star_ms"], st["t_splay_ms"], st["t_out_ms"])
    return _STATE["times"]


def get_tetra():
    """getTetraFromGpu: (tets, opp) in Morton index space (the reordered points)."""
    if _STATE["result"] is None:
        raise RuntimeError("gdel_compute not called")
    tets, opp, _ = _STATE["result"]
    return tets, opp


def gdel_deinit():
    if _STATE.get("tri") is not None:
        _STATE["tri"].close()
    _STATE.update(tri=None, points=None, result=None, times=None)
