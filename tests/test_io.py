"""Point and mesh files through the C ABI (host only) against the restatement in oracle/io_ref.py
and closed forms.  The reference ships no file fixtures for these formats; its readPoints /
_scalePoint / writeToFile code (GD/Main/Application.cpp:475-585, GD/Common/Geometry.cu:395-466)
is what the restatement follows."""
import numpy as np
import pytest

import gstar4d_amd as g
from oracle import io_ref


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_read_asc_dedup_and_power_of_two_scaling(tmp_path):
    # range 4 (exact in float), so scaled = (G - 3) * (v - min) / 4 + 1 exactly
    text = "0 0 0\n1 2 3\n4 4 4\n1 2 3\n0.5 0.25 2\n0 0 0\n"
    path = _write(tmp_path, "a.asc", text)
    pts, sc, info = g.read_asc(path, 64)
    assert info["num_read"] == 6 and info["num_duplicates"] == 2
    assert info["min_val"] == 0.0 and info["max_val"] == 4.0
    want = np.array([[0, 0, 0], [1, 2, 3], [4, 4, 4], [0.5, 0.25, 2]], np.float64)
    assert (pts == want).all()  # first occurrences, in file order
    assert (sc == 61.0 * want / 4.0 + 1.0).all()
    assert sc.min() == 1.0 and sc.max() == 62.0  # [1, G - 2]


def test_read_asc_matches_restatement(tmp_path):
    rng = np.random.default_rng(3)
    v = np.round(rng.uniform(-7.5, 20.0, size=(500, 3)), 3)
    v[10] = v[3]  # duplicates
    v[77] = v[3]
    toks = ["%.3f" % c for c in v.ravel()]
    toks[35] = "7e-1"
    text = " ".join(toks[:700]) + "\n\t" + "\n".join(toks[700:]) + "\n"
    path = _write(tmp_path, "b.asc", text)
    pts, sc, info = g.read_asc(path, 512)
    rp, rs, rinfo = io_ref.read_asc(text, 512)
    assert info == pytest.approx(rinfo)
    assert info["num_duplicates"] == 2
    assert (pts == rp).all()
    assert (sc == rs).all()  # bit-identical float/double mix


def test_read_asc_tokens_and_partial_point(tmp_path):
    # a non-numeric token reads as 0; a trailing partial point is ignored
    text = "1 2 3 abc 5 6 7 8"
    path = _write(tmp_path, "c.asc", text)
    pts, sc, info = g.read_asc(path, 16)
    rp, rs, rinfo = io_ref.read_asc(text, 16)
    assert (pts == np.array([[1, 2, 3], [0, 5, 6]])).all()
    assert (pts == rp).all() and (sc == rs).all()
    assert info["num_read"] == 2 and info["min_val"] == 0.0 and info["max_val"] == 8.0


def test_read_asc_errors(tmp_path):
    with pytest.raises(g.G4DError):
        g.read_asc(str(tmp_path / "missing.asc"), 64)
    # 1.0000000001 reads as the float 1.0, so the second point duplicates the first
    path = _write(tmp_path, "d.asc", "1 1 1\n1.0000000001 1 1\n2 2 2\n")
    pts, _, info = g.read_asc(path, 64)
    assert pts.shape == (2, 3) and info["num_duplicates"] == 1


def test_write_asc_and_ply_text(tmp_path):
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0],
                    [0.123456789, 1e-7, 12345678.0]])
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int32)
    a = tmp_path / "p.asc"
    g.write_asc(str(a), xyz)
    assert a.read_text() == io_ref.asc_text(xyz)
    assert a.read_text().splitlines()[4] == "0.123457 1e-07 1.23457e+07"
    p = tmp_path / "m.ply"
    g.write_ply(str(p), xyz, tets)
    txt = p.read_text()
    assert txt == io_ref.ply_text(xyz, tets)
    lines = txt.splitlines()
    assert lines[:9] == ["ply", "format ascii 1.0", "element vertex 5", "property float x", "property float y",
                         "property float z", "element face 6", "property list uchar int vertex_index", "end_header"]
    assert lines[14:17] == ["3 0 1 2 ", "3 0 1 3 ", "3 0 2 3 "]


def test_asc_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    x = np.round(rng.random((200, 3)), 4)
    a = tmp_path / "r.asc"
    g.write_asc(str(a), x)
    pts, _, info = g.read_asc(str(a), 128)
    assert info["num_read"] == 200
    assert np.allclose(pts, x.astype(np.float32), rtol=0, atol=0)
