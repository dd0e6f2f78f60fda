"""Seeded synthetic meshes and the five benchmark workloads.

No trimesh: the icosphere is an icosahedron subdivided ``s`` times with edge
midpoints projected to the unit sphere (10*4^s + 2 vertices, 20*4^s faces,
outward orientation, children of a face stored consecutively so face tiles
are spatially compact).  Shapes are radial displacements by real orthonormal
spherical harmonics.

Seed convention: shape seed = 1000 * config + k; derived streams use
seed + 1 (oracle frequency subsample), + 2 (finite-difference vertices),
+ 3 (noise), + 4 (upstream gradient).  All draws use
``numpy.random.Generator(PCG64(seed))``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import lpmv

from .render import compute_box_size


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def icosphere(subdivisions: int):
    """Unit icosphere: (verts float64 [10*4^s+2, 3], faces int64 [20*4^s, 3])."""
    v, f = icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(subdivisions):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            j = cache.get(key)
            if j is None:
                p = (np.asarray(verts[a]) + np.asarray(verts[b])) * 0.5
                p /= np.linalg.norm(p)
                verts.append(tuple(p))
                j = len(verts) - 1
                cache[key] = j
            return j

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = np.asarray(new_faces, dtype=np.int64)
    return np.asarray(verts, dtype=np.float64), np.asarray(f, dtype=np.int64)


def real_sph_harm(l: int, m: int, theta, phi):
    """Orthonormal real spherical harmonic Y_lm (theta polar, phi azimuth)."""
    am = abs(m)
    k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    # lpmv includes the Condon-Shortley phase (-1)^m; remove it
    p = lpmv(am, l, np.cos(theta)) * (-1) ** am
    if m == 0:
        return k * p
    if m > 0:
        return math.sqrt(2.0) * k * p * np.cos(am * phi)
    return math.sqrt(2.0) * k * p * np.sin(am * phi)


def sh_radius(unit_verts, coeffs, lmax: int):
    """1 + sum_{l<=lmax, |m|<=l} c[l^2 + l + m] Y_lm at the directions of unit_verts."""
    x, y, z = unit_verts[:, 0], unit_verts[:, 1], unit_verts[:, 2]
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    r = np.ones(len(unit_verts))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            r += coeffs[l * l + l + m] * real_sph_harm(l, m, theta, phi)
    return r


def sh_shape(subdivisions: int, lmax: int, coeff_std: float, seed: int, radius: float = 1.0):
    """Icosphere radially displaced by seeded real SH: r = R (1 + sum c_lm Y_lm)."""
    u, f = icosphere(subdivisions)
    c = rng(seed).normal(0.0, coeff_std, size=(lmax + 1) ** 2)
    r = radius * sh_radius(u, c, lmax)
    return u * r[:, None], f


def needle_edges(faces: np.ndarray, count: int, seed: int):
    """``count`` vertex-disjoint edges, each with its two adjacent faces."""
    edge_faces = {}
    for t, (a, b, c) in enumerate(faces):
        for u, v in ((a, b), (b, c), (c, a)):
            edge_faces.setdefault((min(u, v), max(u, v)), []).append(t)
    edges = [e for e, fs in edge_faces.items() if len(fs) == 2]
    order = rng(seed).permutation(len(edges))
    used, picked = set(), []
    for i in order:
        e = edges[i]
        touched = {int(x) for t in edge_faces[e] for x in faces[t]}
        if touched & used:
            continue
        used |= touched
        picked.append(e)
        if len(picked) == count:
            break
    needle_faces = sorted(t for e in picked for t in edge_faces[e])
    return picked, np.asarray(needle_faces, dtype=np.int64)


@dataclass
class Workload:
    name: str
    verts: np.ndarray          # float32 [nv, 3]
    faces: np.ndarray          # int64 [nf, 3]
    coeff: np.ndarray          # float32 [nf]
    box_size: np.ndarray       # float64 [3, 2]
    box_shape: tuple
    sigma: np.ndarray          # float32 [3, 3]
    thresh: float
    seed: int
    extra: dict = field(default_factory=dict)


def config1(k: int = 0) -> Workload:
    """Icosphere s3 (642 v / 1280 f), R = 0.3, jitter N(0, 0.005); 64^3; Sigma = 1e-3 I; thresh 1e-5."""
    seed = 1000 * 1 + k
    u, f = icosphere(3)
    v = 0.3 * u + rng(seed).normal(0.0, 0.005, size=u.shape)
    v = v.astype(np.float32)
    return Workload("cfg1_ico3_64", v, f, np.ones(len(f), np.float32),
                    compute_box_size(v, offset=0.2), (64, 64, 64),
                    (1e-3 * np.eye(3)).astype(np.float32), 1e-5, seed)


def config2(k: int = 0, radius: float = 1.0) -> Workload:
    """Icosphere s5 (10242 v / 20480 f), SH l <= 6, c ~ N(0, 0.03), R = 1; 200^3; Sigma = 1e-3 I; thresh 1e-5."""
    seed = 2000 + k
    v, f = sh_shape(5, 6, 0.03, seed, radius)
    v = v.astype(np.float32)
    return Workload("cfg2_ico5_sh6_200", v, f, np.ones(len(f), np.float32),
                    compute_box_size(v, offset=0.2), (200, 200, 200),
                    (1e-3 * np.eye(3)).astype(np.float32), 1e-5, seed)


def config3(k: int = 0) -> Workload:
    """Config 2's shape family plus 410 near-degenerate faces: 205 vertex-disjoint
    edges shrunk to 1e-4 of their length (one endpoint moved towards the
    other), so the two faces on each edge become needles.  Sigma =
    diag(1e-3, 1e-3, 4e-3); thresh 1e-7; box 256 x 256 x 128."""
    seed = 3000 + k
    v, f = sh_shape(5, 6, 0.03, seed, 1.0)
    edges, needles = needle_edges(f, 205, seed)
    for a, b in edges:
        v[b] = v[a] + 1e-4 * (v[b] - v[a])
    v = v.astype(np.float32)
    return Workload("cfg3_ico5_needles_256x256x128", v, f, np.ones(len(f), np.float32),
                    compute_box_size(v, offset=0.2), (256, 256, 128),
                    np.diag([1e-3, 1e-3, 4e-3]).astype(np.float32), 1e-7, seed,
                    extra={"needle_faces": needles, "needle_edges": np.asarray(edges)})


def config4(k: int = 0, subdivisions: int = 4, box: int = 128) -> Workload:
    """Fit workload: target = icosphere s4 with SH l <= 8, c ~ N(0, 0.05); the
    initial mesh (in extra['init_verts']) is the unit icosphere s4; 128^3,
    Sigma = 1e-3 I, thresh 1e-5."""
    seed = 4000 + k
    v, f = sh_shape(subdivisions, 8, 0.05, seed, 1.0)
    u, _ = icosphere(subdivisions)
    v = v.astype(np.float32)
    box_size = compute_box_size(np.concatenate([v, u]), offset=0.2)
    return Workload(f"cfg4_fit_ico{subdivisions}_{box}", v, f, np.ones(len(f), np.float32),
                    box_size, (box, box, box), (1e-3 * np.eye(3)).astype(np.float32), 1e-5, seed,
                    extra={"init_verts": u.astype(np.float32)})


def doublet(subdivisions: int = 3, seed: int = 6000):
    """A two-cell multi-material mesh in the ``.rec`` labelling: an icosphere
    (jittered by N(0, 0.003)) cut by the plane z = 0 into cell 1 (face
    centroids above) and cell 2 (below), whose exterior faces carry labels
    (1, 0) / (2, 0), plus an interface disc fanned from a centre vertex to the
    loop of edges where the two halves meet, labelled (1, 2).  Returns
    (verts [nv, 3] float64, faces_mm [nf, 5] int64)."""
    u, f = icosphere(subdivisions)
    v = u + rng(seed).normal(0.0, 0.003, size=u.shape)
    top = v[f].mean(1)[:, 2] > 0.0
    labels = np.where(top, 1, 2)
    # boundary loop: directed edges of the top faces whose reverse is a bottom-face edge
    top_edges = {(a, b) for tri in f[top] for a, b in ((tri[0], tri[1]), (tri[1], tri[2]),
                                                        (tri[2], tri[0]))}
    bot_edges = {(a, b) for tri in f[~top] for a, b in ((tri[0], tri[1]), (tri[1], tri[2]),
                                                         (tri[2], tri[0]))}
    loop = sorted((a, b) for a, b in top_edges if (b, a) in bot_edges)
    centre = len(v)
    loop_verts = np.unique(np.array(loop).ravel())
    v = np.vstack([v, [[0.0, 0.0, float(v[loop_verts, 2].mean())]]])
    faces = [list(tri) + [int(lab), 0] for tri, lab in zip(f, labels)]
    faces += [[a, b, centre, 1, 2] for a, b in loop]
    return v, np.array(faces, dtype=np.int64)


def noisy_image(img: np.ndarray, seed: int, photons: float = 500.0, sigma: float = 0.01):
    """Poisson(photons * img) / photons + N(0, sigma^2) (config 4's noise model)."""
    g = rng(seed)
    return (g.poisson(photons * np.clip(img, 0, None)) / photons
            + g.normal(0.0, sigma, size=img.shape)).astype(np.float32)
