"""The oracle against the committed golden vectors of tests/golden/
(mpmath quadrature of single faces, exact unit-cube transforms; written by
tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def cplx(pair):
    a = np.asarray(pair, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


FACES = load("single_faces.json")


@pytest.mark.parametrize("case", FACES, ids=[c["case"] for c in FACES])
def test_oracle_face_vs_quadrature_vectors(case):
    F, dF = O.face_accurate(*case["phases"])
    want = cplx(case["value"])
    dwant = cplx(case["dvalue_dphase"])
    assert abs(2 * F - want) <= 1e-11 * max(abs(want), 1e-3)
    assert np.max(np.abs(2 * dF - dwant)) <= 1e-9 * max(np.max(np.abs(dwant)), 1e-3)


@pytest.mark.parametrize("case", FACES, ids=[c["case"] for c in FACES])
def test_oracle_one_face_mesh_vs_quadrature_vectors(case):
    """The oracle's mesh path (weights, phases from FP32 inputs) on the
    one-face mesh: value and vertex gradient."""
    V = np.array(case["verts"])
    xi = np.array([case["xi"]])
    f = np.array([[0, 1, 2]])
    got = O.acc_forward(V, f, np.ones(1), xi, threads=1)
    want = cplx(case["value"])
    assert abs(got[0] - want) <= 1e-11 * max(abs(want), 1e-3)
    g = np.array([0.6 - 0.8j])
    gv = O.acc_backward(V, f, np.ones(1), xi, g, threads=1)
    # d/dv_j Re(conj(g) value) = Re(conj(g) dvalue/dx_j) xi
    gw = np.real(np.conj(g[0]) * cplx(case["dvalue_dphase"]))[:, None] * xi
    assert np.max(np.abs(gv - gw)) <= 1e-9 * max(np.max(np.abs(gw)), 1e-3)


def test_oracle_cube_vs_exact_vectors():
    c = load("cube_surface.json")
    V, f, xi = np.array(c["verts"], float), np.array(c["faces"]), np.array(c["xi"])
    surf, solid = cplx(c["surface"]), cplx(c["solid"])
    got = O.acc_forward(V, f, np.ones(len(f)), xi, threads=2)
    assert np.max(np.abs(got - surf)) <= 1e-12
    af = O.acc_forward_af(V, f, np.ones(len(f)), 0.7, xi, threads=2)
    want = surf - 0.7 * solid
    want[np.all(xi == 0, axis=1)] = 1.0 - 0.7
    assert np.max(np.abs(af - want)) <= 1e-9


def test_cube_fixture_matches_numpy_closed_form():
    """The committed cube vectors against the same closed form in numpy
    (double precision, the 1-D integral written as exp(-ik/2) sin(k/2)/(k/2)
    to avoid cancellation)."""
    c = load("cube_surface.json")
    xi = np.array(c["xi"])

    def I(k):
        return 1.0 if k == 0 else np.exp(-0.5j * k) * np.sin(0.5 * k) / (0.5 * k)

    surf = np.array([(I(a) * I(b) * (1 + np.exp(-1j * z)) + I(b) * I(z) * (1 + np.exp(-1j * a))
                      + I(z) * I(a) * (1 + np.exp(-1j * b))) / 6.0 for a, b, z in xi])
    solid = np.array([I(a) * I(b) * I(z) for a, b, z in xi])
    assert np.max(np.abs(surf - cplx(c["surface"]))) <= 1e-12
    assert np.max(np.abs(solid - cplx(c["solid"]))) <= 1e-12
