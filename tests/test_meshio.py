"""Mesh files: the multi-material .rec format, OBJ, and the interface-label
-> Faces_coeff rule of the reference's multi-material example."""
import os

import numpy as np
import pytest

from deltamic_amd import synthetic as syn
from deltamic_amd.meshio import multimaterial_coefficients, read_obj, read_rec, write_rec

REF_MESH = "/root/reference/data/mesh"


def test_rec_round_trip_and_coefficients(tmp_path):
    v, f = syn.doublet()
    p = tmp_path / "doublet.rec"
    write_rec(p, v, f)
    v2, f2 = read_rec(p)
    assert np.array_equal(v2, v) and np.array_equal(f2, f)
    c = multimaterial_coefficients(f)
    assert c.dtype == np.float32
    # exterior faces (one label 0) weigh 1, cell-cell faces 2
    assert np.array_equal(c == 1.0, (f[:, 3] == 0) | (f[:, 4] == 0))
    assert set(np.unique(c)) == {1.0, 2.0}
    # the interface loop edges are shared by three faces (non-manifold)
    edges = {}
    for tri in f[:, :3]:
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            key = (min(a, b), max(a, b))
            edges[key] = edges.get(key, 0) + 1
    assert sorted(set(edges.values())) == [2, 3]


def test_rec_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.rec"
    p.write_text("2\n0 0 0\n1 1 1\n1\n0 1 2 0 1\n")
    with pytest.raises(ValueError):
        read_rec(p)
    p.write_text("1\n0 0 0\n0\n7\n")
    with pytest.raises(ValueError):
        read_rec(p)


def test_obj_reader_forms(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("# quad and triangle\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
                 "f 1//1 2//1 3//1 4//1\nf -4/1 -2/1 -1/1\n")
    v, f = read_obj(p)
    assert v.shape == (4, 3)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 3]]


@pytest.mark.skipif(not os.path.isdir(REF_MESH), reason="reference data files not present")
def test_reference_mesh_files_parse():
    """The reference's example meshes (inputs only): counts as listed in its data."""
    v, f = read_rec(os.path.join(REF_MESH, "Sample01_007.rec"))
    assert v.shape == (1996, 3) and f.shape == (4055, 5)
    c = multimaterial_coefficients(f)
    assert (c == 2.0).sum() > 0 and (c == 1.0).sum() > 0
    v, f = read_rec(os.path.join(REF_MESH, "cell4.rec"))
    assert v.shape == (8162, 3) and f.shape == (16541, 5)
    v, f = read_obj(os.path.join(REF_MESH, "cow.obj"))
    assert v.shape == (2930, 3) and f.shape == (5856, 3)
