"""CPU restatement of the reference's point-file reader, grid scaling and file writers (test
infrastructure only: imported by tests/, never by the product library).

Follows:
  Application::readPoints      GD/Main/Application.cpp:475-566
  Application::_scalePoint     GD/Main/Application.cpp:568-585
  Input-Points.asc dump        GD/Main/Application.cpp:231-243
  TetraMesh::writeToFile       GD/Common/Geometry.cu:395-466
Tokens are parsed to float32 through float64 (numpy), which equals a direct decimal-to-float
conversion for the short decimals the tests use.
"""
import re

import numpy as np

_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _token_value(tok):
    # stream extraction of a float: the longest numeric prefix, 0 when there is none
    m = _NUM.match(tok)
    return np.float32(float(m.group(0))) if m else np.float32(0.0)


def read_asc(text, grid_size):
    vals = [_token_value(t) for t in text.split()]
    vmin, vmax = np.float32(999.0), np.float32(-999.0)
    pts, seen = [], set()
    cur = []
    n_read = 0
    for v in vals:
        cur.append(v)
        vmin = min(vmin, v)
        vmax = max(vmax, v)
        if len(cur) == 3:
            n_read += 1
            key = tuple(float(c) for c in cur)
            if key not in seen:
                seen.add(key)
                pts.append(key)
            cur = []
    pts = np.array(pts, np.float64).reshape(-1, 3)
    rng = np.float64(np.float32(vmax - vmin))           # float difference, used as a double
    scaled = (np.float64(grid_size) - 3.0) * (pts - np.float64(vmin)) / rng + 1.0
    return pts, scaled, {"num_read": n_read, "num_duplicates": n_read - pts.shape[0],
                         "min_val": float(vmin), "max_val": float(vmax)}


def _g(x):
    return "%g" % x


def asc_text(xyz):
    return "".join("%s %s %s\n" % (_g(p[0]), _g(p[1]), _g(p[2])) for p in np.asarray(xyz, np.float64))


def ply_text(xyz, tets):
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    tets = np.asarray(tets, np.int64).reshape(-1, 4)
    out = ["ply", "format ascii 1.0", "element vertex %d" % xyz.shape[0], "property float x",
           "property float y", "property float z", "element face %d" % (3 * tets.shape[0]),
           "property list uchar int vertex_index", "end_header"]
    lines = "\n".join(out) + "\n"
    lines += "".join("%s %s %s \n" % (_g(p[0]), _g(p[1]), _g(p[2])) for p in xyz)
    for t in tets:
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3)):
            lines += "3 %d %d %d \n" % (t[f[0]], t[f[1]], t[f[2]])
    return lines
