"""Mesh files for the fit: the multi-material ``.rec`` format and OBJ.

The reference reads ``.rec`` files with an external package
(``dw3d.open_mesh_multitracker``, Examples/.ipynb_checkpoints/
Example_multimaterial-checkpoint.ipynb cell 0) that is not part of it; the
format, as its data files show (data/mesh/cell4.rec, Sample01_007.rec), is

    nv
    x y z            (nv lines)
    nf
    v0 v1 v2 a b     (nf lines: a triangle and the labels of the two regions
                      it separates; label 0 is the exterior)

and the multi-material example turns the labels into ``Faces_coeff``: 1 for
a face on the exterior (either label 0), 2 for a face between two cells
(notebook cell 1).  This module restates both with numpy; the coefficients
then go to ``Fourier3dMesh.forward`` unchanged.
"""
from __future__ import annotations

import numpy as np


def read_rec(path):
    """(verts [nv, 3] float64, faces_mm [nf, 5] int64: v0 v1 v2 label_a label_b)."""
    with open(path) as fh:
        tokens = fh.read().split()
    pos = 0
    nv = int(tokens[pos]); pos += 1
    verts = np.array(tokens[pos:pos + 3 * nv], dtype=np.float64).reshape(nv, 3)
    pos += 3 * nv
    nf = int(tokens[pos]); pos += 1
    faces = np.array(tokens[pos:pos + 5 * nf], dtype=np.int64).reshape(nf, 5)
    pos += 5 * nf
    if pos != len(tokens):
        raise ValueError(f"{path}: {len(tokens) - pos} unexpected trailing tokens")
    if nf and (faces[:, :3].min() < 0 or faces[:, :3].max() >= nv):
        raise ValueError(f"{path}: face index outside [0, {nv})")
    return verts, faces


def write_rec(path, verts, faces_mm):
    """Writes the format read_rec reads (coordinates with 17 significant digits)."""
    verts = np.asarray(verts, dtype=np.float64)
    faces_mm = np.asarray(faces_mm, dtype=np.int64)
    with open(path, "w") as fh:
        fh.write(f"{len(verts)}\n")
        for x, y, z in verts:
            fh.write(f"{x:.17g} {y:.17g} {z:.17g}\n")
        fh.write(f"{len(faces_mm)}\n")
        for row in faces_mm:
            fh.write(" ".join(str(int(v)) for v in row) + "\n")


def multimaterial_coefficients(faces_mm) -> np.ndarray:
    """Faces_coeff of a multi-material mesh: 1 where a face borders the
    exterior (label 0 on either side), 2 between two cells."""
    f = np.asarray(faces_mm)
    exterior = (f[:, 3] == 0) | (f[:, 4] == 0)
    return np.where(exterior, 1.0, 2.0).astype(np.float32)


def read_obj(path):
    """(verts [nv, 3] float64, faces [nf, 3] int64) of a triangle OBJ file
    ('v x y z' and 'f a b c' records; 'a/b/c' index forms and negative
    indices accepted; polygons are fanned into triangles)."""
    verts, faces = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(p) for p in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for p in parts[1:]:
                    i = int(p.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                for k in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[k], idx[k + 1]])
    return np.array(verts, dtype=np.float64).reshape(-1, 3), np.array(faces, dtype=np.int64).reshape(-1, 3)
