"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is independent of the oracle's own formulas:
  * mpmath quadrature of the face integral over a triangle (every branch,
    including coincident and nearly coincident phases),
  * the closed-form transform of a unit-cube surface (separable 1-D integrals),
  * invariants: corner derivatives sum to -iF, translation phase, Hermitian
    symmetry, invariance under uniform coefficient scaling and rotation,
  * central finite differences of the FP64 forward for the adjoint,
  * the sphere limit sin(|xi|R)/(|xi|R).
The reference-semantics restatement (orc_ref_*) is checked against the
accurate oracle where the reference's formulas are well conditioned.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O
from deltamic_amd import synthetic as syn

mp.mp.dps = 40


def quad_face(xa, p, q, weight=None):
    """int_{s,t>=0, s+t<=1} w(s,t) exp(-i(xa + s p + t q)) ds dt by mpmath.

    The t-integral is done in closed form, the s-integral numerically.
    weight: None -> 1, 's' -> -i s (dF/dx_b), 't' -> -i t (dF/dx_c)."""
    xa, p, q = mp.mpf(xa), mp.mpf(p), mp.mpf(q)
    I = mp.mpc(0, 1)

    def inner(s):  # int_0^{1-s} w exp(-i t q) dt
        L = 1 - s
        if weight == 't':
            if q == 0:
                return -I * L * L / 2
            # int_0^L -i t e^{-itq} dt
            e = mp.exp(-I * q * L)
            return -I * ((1 - e * (1 + I * q * L)) / (-(q ** 2)))
        if q == 0:
            base = L
        else:
            base = (1 - mp.exp(-I * q * L)) / (I * q)
        return base * (-I * s if weight == 's' else 1)

    f = lambda s: mp.exp(-I * (xa + s * p)) * inner(s)
    pts = np.linspace(0, 1, 2 + int(min(64, abs(float(p)) + abs(float(q))))).tolist()
    v = mp.quad(f, [mp.mpf(x) for x in pts])
    return complex(v)


CASES = [
    (0.3, 1.7, -0.9),        # generic
    (500.25, 2.5, 4.0),      # large absolute phase
    (-3.0, 40.0, -25.0),     # large phase gaps
    (1.0, 1e-3, 2e-3),       # all close (Taylor)
    (1.0, 1e-9, -2e-9),      # all nearly coincident
    (2.0, 0.0, 3.0),         # a == b exactly
    (2.0, 1e-12, 3.0),       # a ~ b, c far (quad branch)
    (2.0, 3.0, 3.0),         # b == c exactly
    (2.0, -1e-7, 1.5),       # near-coincident pair
    (0.0, 0.0, 0.0),         # all equal
    (7.0, 0.5, -0.5),        # symmetric about a
    (1.0, 0.9, 1.1),         # taylor boundary region
]


@pytest.mark.parametrize("xa,p,q", CASES)
def test_face_value_and_derivatives_vs_quadrature(xa, p, q):
    F, dF = O.face_accurate(xa, xa + p, xa + q)
    Fq = quad_face(xa, p, q)
    dbq = quad_face(xa, p, q, 's')
    dcq = quad_face(xa, p, q, 't')
    daq = -1j * Fq - dbq - dcq
    tol = 1e-11
    assert abs(F - Fq) <= tol * max(abs(Fq), 1e-3)
    for got, want in zip(dF, (daq, dbq, dcq)):
        assert abs(got - want) <= 1e-9 * max(abs(want), 1e-3), (got, want)


@pytest.mark.parametrize("xa,p,q", CASES)
def test_corner_derivatives_sum_to_minus_iF(xa, p, q):
    F, dF = O.face_accurate(xa, xa + p, xa + q)
    assert abs(dF.sum() + 1j * F) <= 1e-12 * max(1.0, abs(F))


def test_face_symmetric_in_its_nodes():
    g = np.random.default_rng(3)
    for _ in range(200):
        x = g.normal(0, 3, 3) + g.uniform(-300, 300)
        F0, d0 = O.face_accurate(*x)
        for perm in ((1, 2, 0), (2, 0, 1), (1, 0, 2)):
            F1, d1 = O.face_accurate(*x[list(perm)])
            assert abs(F1 - F0) <= 1e-12
            assert np.allclose(d1, d0[list(perm)], atol=1e-11, rtol=0)


def test_reference_semantics_face_matches_where_well_conditioned():
    g = np.random.default_rng(4)
    for _ in range(300):
        xa = g.uniform(-100, 100)
        p, q = g.uniform(0.5, 6, 2) * g.choice([-1, 1], 2)
        if abs(p - q) < 0.5:
            continue
        F, dF = O.face_accurate(xa, xa + p, xa + q)
        Fr, dFr = O.face_ref(xa, xa + p, xa + q, np.float64)
        assert abs(Fr - F) <= 1e-11
        assert np.allclose(dFr, dF, atol=1e-10, rtol=0)


def _cube():
    v = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], dtype=np.float64)
    idx = {tuple(p): i for i, p in enumerate(v.astype(int))}
    quads = [  # outward-facing quads, counter-clockwise seen from outside
        [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)],  # z = 0, normal -z
        [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],  # z = 1
        [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)],  # y = 0
        [(0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)],  # y = 1
        [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)],  # x = 0
        [(1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)],  # x = 1
    ]
    faces = []
    for qd in quads:
        a, b, c, d = (idx[p] for p in qd)
        faces += [(a, b, c), (a, c, d)]
    return v, np.array(faces, dtype=np.int64)


def _cube_ft(xi):
    def I(k):  # int_0^1 exp(-ikx) dx = exp(-ik/2) sin(k/2)/(k/2), without cancellation
        return 1.0 if k == 0 else np.exp(-0.5j * k) * np.sin(0.5 * k) / (0.5 * k)
    out = []
    for kx, ky, kz in xi:
        s = (I(kx) * I(ky) * (1 + np.exp(-1j * kz)) + I(ky) * I(kz) * (1 + np.exp(-1j * kx))
             + I(kz) * I(kx) * (1 + np.exp(-1j * ky)))
        out.append(s / 6.0)
    return np.array(out)


def test_cube_surface_closed_form():
    v, f = _cube()
    g = np.random.default_rng(5)
    xi = np.concatenate([
        g.normal(0, 6, (200, 3)),
        np.array([[0, 0, 0], [3.0, 0, 0], [0, -2.0, 0], [0, 0, 7.5], [1.0, 1.0, 0], [2 * np.pi, 0, 0],
                  [1e-9, 2.0, 0], [40.0, -30.0, 0.5]]),
    ])
    want = _cube_ft(xi)
    got = O.acc_forward(v, f, np.ones(len(f)), xi, threads=2, fp32_inputs=False)
    assert np.max(np.abs(got - want)) <= 1e-12
    # the reference-semantics FP64 path agrees away from its sqrt(eps) branches
    ref = O.ref_forward(v, f, np.ones(len(f)), xi, np.float64, threads=2)
    assert np.max(np.abs(ref - want)) <= 1e-6


def _cube_solid_ft(xi):
    """int over [0,1]^3 of exp(-i xi.x) dx: a product of 1-D integrals."""
    def I(k):
        return 1.0 if k == 0 else np.exp(-0.5j * k) * np.sin(0.5 * k) / (0.5 * k)
    return np.array([I(a) * I(b) * I(c) for a, b, c in xi])


def test_autofluorescence_cube_closed_form():
    """Autofluorescence variant (render.py:228-318) on the unit cube: the
    interior term is -regint times the solid cube's transform over its
    volume (divergence theorem), for either orientation of the surface, and
    its value at xi = 0 is the limit -regint."""
    v, f = _cube()
    g = np.random.default_rng(6)
    xi = np.concatenate([g.normal(0, 6, (150, 3)),
                         np.array([[3.0, 0, 0], [0, 0, 7.5], [1e-7, 2e-7, -1e-7], [40.0, -30.0, 0.5]])])
    for faces in (f, f[:, ::-1].copy()):
        surf = O.acc_forward(v, faces, np.ones(len(faces)), xi, threads=2, fp32_inputs=False)
        af = O.acc_forward_af(v, faces, np.ones(len(faces)), 0.7, xi, threads=2, fp32_inputs=False)
        assert np.max(np.abs((af - surf) - (-0.7 * _cube_solid_ft(xi)))) <= 1e-9
        at0 = O.acc_forward_af(v, faces, np.ones(len(faces)), 0.7, np.zeros((1, 3)), fp32_inputs=False)
        assert abs(at0[0] - (1.0 - 0.7)) <= 1e-12


def _small_mesh(seed=11, sub=2):
    u, f = syn.icosphere(sub)
    g = syn.rng(seed)
    v = 0.8 * u + g.normal(0, 0.02, u.shape) + np.array([1.2, 1.1, 1.3])
    return v, f


def _freqs(seed, n, scale=6.0):
    return syn.rng(seed).normal(0, scale, (n, 3))


def test_invariants_translation_hermitian_scaling_rotation():
    v, f = _small_mesh()
    c = syn.rng(12).uniform(0.5, 2.0, len(f))
    xi = _freqs(13, 64)
    F0 = O.acc_forward(v, f, c, xi, threads=4, fp32_inputs=False)
    t = np.array([0.3, -0.7, 0.25])
    Ft = O.acc_forward(v + t, f, c, xi, threads=4, fp32_inputs=False)
    assert np.max(np.abs(Ft - np.exp(-1j * xi @ t) * F0)) <= 1e-11
    Fm = O.acc_forward(v, f, c, -xi, threads=4, fp32_inputs=False)
    assert np.max(np.abs(Fm - np.conj(F0))) <= 1e-12
    Fs = O.acc_forward(v, f, 3.7 * c, xi, threads=4, fp32_inputs=False)
    assert np.max(np.abs(Fs - F0)) <= 1e-12
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Fr = O.acc_forward(v @ R.T, f, c, xi @ R.T, threads=4, fp32_inputs=False)
    assert np.max(np.abs(Fr - F0)) <= 1e-11
    assert abs(O.acc_forward(v, f, c, np.zeros((1, 3)), fp32_inputs=False)[0] - 1.0) <= 1e-14


def test_sphere_limit():
    u, f = syn.icosphere(5)
    R, ctr = 0.9, np.array([0.1, -0.2, 0.3])
    xi = _freqs(14, 32, scale=2.0)
    got = O.acc_forward(R * u + ctr, f, np.ones(len(f)), xi, threads=8, fp32_inputs=False)
    k = np.linalg.norm(xi, axis=1) * R
    want = np.sin(k) / k * np.exp(-1j * xi @ ctr)
    assert np.max(np.abs(got - want)) <= 2e-3


def test_adjoint_vs_central_finite_differences():
    """64 seeded vertices (seed + 2), h = 1e-3 in FP64, non-uniform coefficients."""
    v, f = _small_mesh(sub=2)
    c = syn.rng(21).uniform(0.5, 2.0, len(f))
    xi = _freqs(22, 300)
    g = syn.rng(24)
    grad = g.standard_normal(len(xi)) + 1j * g.standard_normal(len(xi))
    an = O.acc_backward(v, f, c, xi, grad, threads=8, fp32_inputs=False)

    def L(V):
        return float(np.sum(np.real(np.conj(grad) * O.acc_forward(V, f, c, xi, threads=8,
                                                                    fp32_inputs=False))))

    verts = np.sort(syn.rng(20 + 2).choice(len(v), 64, replace=False))
    h = 1e-3
    fd = np.zeros((64, 3))
    for i, vi in enumerate(verts):
        for d in range(3):
            vp = v.copy(); vp[vi, d] += h
            vm = v.copy(); vm[vi, d] -= h
            vp2 = v.copy(); vp2[vi, d] += 2 * h
            vm2 = v.copy(); vm2[vi, d] -= 2 * h
            # 4th-order central difference
            fd[i, d] = (-L(vp2) + 8 * L(vp) - 8 * L(vm) + L(vm2)) / (12 * h)
    err = np.linalg.norm(fd - an[verts]) / np.linalg.norm(an[verts])
    assert err <= 1e-6, err


def test_adjoint_face_partials_identity():
    """P_c = T - P_a - P_b reproduces the per-corner sums computed directly."""
    v, f = _small_mesh(sub=1)
    xi = _freqs(31, 40)
    g = syn.rng(32)
    grad = g.standard_normal(len(xi)) + 1j * g.standard_normal(len(xi))
    _, part = O.acc_backward(v, f, np.ones(len(f)), xi, grad, threads=2, return_face=True,
                             fp32_inputs=False)
    for t in range(0, len(f), 7):
        pc = np.zeros(3)
        for m, x in enumerate(xi):
            xs = [float(v[j] @ x) for j in f[t]]
            _, dF = O.face_accurate(*xs)
            pc += np.real(np.conj(grad[m]) * dF[2]) * x
        assert np.allclose(part[t, 6:9] - part[t, 0:3] - part[t, 3:6], pc, atol=1e-10)


def test_reference_semantics_adjoint_f64_agrees_on_well_conditioned_mesh():
    v, f = _small_mesh(sub=2)
    c = np.ones(len(f))
    xi = _freqs(41, 200, scale=30.0)
    g = syn.rng(42)
    grad = g.standard_normal(len(xi)) + 1j * g.standard_normal(len(xi))
    ft = O.acc_forward(v, f, c, xi, threads=4, fp32_inputs=False)
    ref_ft = O.ref_forward(v, f, c, xi, np.float64, threads=4)
    assert O.rel_l2(ref_ft, ft) <= 1e-9
    an = O.acc_backward(v, f, c, xi, grad, threads=4, fp32_inputs=False)
    rb = O.ref_backward(v, f, c, xi, grad, ft, np.float64, threads=4)
    assert O.rel_l2(rb, an) <= 1e-7


def test_band_restatement_order_and_threshold():
    f = np.array([[[0.5, 2e-5], [1e-5, 0.0]], [[3e-5, 1.0], [1e-6, 0.2]]], dtype=np.float32)
    idx, xi = O.band(f, np.float32(1e-5), [10.0, 20.0], [1.0, 2.0], [0.1, 0.2])
    assert idx.tolist() == [0, 1, 4, 5, 7]  # strict '>', row-major, 1e-5 itself excluded
    assert xi.tolist()[2] == [20.0, 1.0, 0.1]
