"""Writes the golden vectors of tests/golden/ from mathematics alone.

    python tests/golden/make_golden.py      (about five minutes; needs mpmath)

The reference holds no golden vectors for the mesh Fourier transform, so the
vectors here come from two computations that share nothing with the oracle
or the HIP kernels:

single_faces.json
    One triangle and one frequency per case, on FP32-exact geometry.  With
    f(x) = -exp(-ix), a face's value in a one-face mesh is
    2 f[x_a, x_b, x_c] (the face weight 2 c A / sum(c A) is 2), and by the
    Hermite-Genocchi formula
        f[x_a, x_b, x_c] = int_{s,t>=0, s+t<=1} exp(-i(x_a + s p + t q)) ds dt,
    p = x_b - x_a, q = x_c - x_a (the reference's fourier_xi_tri_f,
    deltamic/cpp/fourier3d_cuda_kernel.cu:239-278, evaluates the same divided
    difference in closed form).  The t-integral is done in closed form and
    the s-integral by mpmath quadrature at 30 digits; d/dx_b and d/dx_c
    weigh the integrand by -i s and -i t, and d/dx_a = -i F - d/dx_b - d/dx_c.
    The phases x_j = xi . v_j are formed exactly from the FP32 inputs.
    The cases cover every branch of the reference (.cu:239-278: all phases
    distinct, one pair coincident, all coincident) and of the HIP kernels
    (Taylor branch for phase gaps below 0.1 rad, the closed form, gaps
    above 1000 rad).  Coordinates are dyadic, so the FP32 edge vectors and
    phase differences the kernels form are exact.

cube_surface.json
    The unit cube [0,1]^3 as 12 triangles: its surface transform over its
    area 6, a sum of products of 1-D integrals
        I(k) = int_0^1 exp(-ikx) dx = (1 - exp(-ik))/(ik),
    and the solid cube's transform I(kx) I(ky) I(kz), which gives the
    autofluorescence variant (deltamic/render.py:228-318) by the divergence
    theorem: surface - regint * solid (limit 1 - regint at xi = 0).
    Frequencies are multiples of 1/4 (exact in FP32), including gaps above
    1000 rad.
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))
I1 = mp.mpc(0, 1)


def _phases(verts, xi):
    return [mp.fsum(mp.mpf(float(xi[d])) * mp.mpf(float(v[d])) for d in range(3)) for v in verts]


def quad_face(xa, p, q, weight=None):
    """int over the unit simplex of w(s,t) exp(-i(xa + s p + t q)) ds dt."""
    def inner(s):  # int_0^{1-s} w(s,t) exp(-i t q) dt
        L = 1 - s
        if weight == "t":
            if q == 0:
                return -I1 * L * L / 2
            e = mp.exp(-I1 * q * L)
            return -I1 * ((1 - e * (1 + I1 * q * L)) / (-(q ** 2)))
        base = L if q == 0 else (1 - mp.exp(-I1 * q * L)) / (I1 * q)
        return base * (-I1 * s if weight == "s" else 1)

    f = lambda s: mp.exp(-I1 * (xa + s * p)) * inner(s)
    n = 2 + int(min(12000, (abs(p) + abs(q)) / 2))
    return mp.quad(f, mp.linspace(0, 1, n))


# (xi, a, b, c): xi a multiple of 1/4 (or 2), vertices dyadic.
FACE_CASES = [
    ("generic", (1.25, -0.75, 2.5), (0.125, 0.5, -0.25), (1.375, 0.25, 0.5), (0.5, 1.5, -0.125)),
    ("large absolute phase", (2.0, 0.0, 0.0), (250.125, 0.0, 0.0), (251.375, 1.0, 0.0), (252.125, 0.0, 1.0)),
    ("large gaps", (4.0, 0.0, 0.0), (-0.75, 0.0, 0.0), (9.25, 1.0, 0.0), (-7.0, 0.0, 1.0)),
    ("taylor: all gaps below 0.1", (2.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.5078125, 1.0, 0.0), (0.484375, 0.0, 1.0)),
    ("all nearly coincident (gaps below 5e-7)", (2.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0000002384185791, 1.0, 0.0),
     (0.9999998807907104, 0.0, 1.0)),
    ("a == b, c far", (2.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (2.5, 0.0, 1.0)),
    ("b == c", (2.0, 0.0, 0.0), (1.0, 0.0, 0.0), (2.5, 1.0, 0.0), (2.5, 0.0, 1.0)),
    ("near pair, third far", (2.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.9999998807907104, 1.0, 0.0), (1.75, 0.0, 1.0)),
    ("all phases zero, xi normal to the face", (0.0, 0.0, 2.0), (0.25, 0.5, 0.0), (1.25, 0.5, 0.0), (0.25, 1.5, 0.0)),
    ("xi = 0", (0.0, 0.0, 0.0), (0.25, 0.5, 0.75), (1.25, 0.5, 0.0), (0.25, 1.5, 0.0)),
    ("symmetric about a", (2.0, 0.0, 0.0), (3.5, 0.0, 0.0), (3.75, 1.0, 0.0), (3.25, 0.0, 1.0)),
    ("gap near the Taylor threshold", (2.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.546875, 1.0, 0.0), (0.5, 0.0, 1.0)),
    ("phase series region |P| < 1", (2.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.625, 1.0, 0.0), (0.0, 0.0, 1.0)),
    ("wide: gaps above 1000 rad", (4.0, 0.0, 0.0), (2.5, 0.0, 0.0), (377.5, 1.0, 0.0), (-172.5, 0.0, 1.0)),
    ("very wide: gaps near 2e4 rad", (16.0, 0.0, 0.0), (0.5, 0.0, 0.0), (800.25, 1.0, 0.0), (-400.5, 0.0, 1.0)),
    ("wide, oblique", (3.25, -2.5, 1.75), (10.125, -3.0, 2.5), (190.5, 40.25, -100.0), (-300.0, 150.5, 60.125)),
]


def single_faces():
    out = []
    for name, xi, a, b, c in FACE_CASES:
        xi32 = np.array(xi, np.float32)
        V32 = np.array([a, b, c], np.float32)
        assert np.array_equal(V32.astype(np.float64), np.array([a, b, c])), name  # exact in FP32
        xa, xb, xc = _phases(V32, xi32)
        p, q = xb - xa, xc - xa
        F = quad_face(xa, p, q)
        dB = quad_face(xa, p, q, "s")
        dC = quad_face(xa, p, q, "t")
        dA = -I1 * F - dB - dC
        out.append({
            "case": name,
            "xi": [float(x) for x in xi32],
            "verts": V32.astype(float).tolist(),
            "phases": [float(xa), float(xb), float(xc)],
            # value of the one-face mesh and its derivatives w.r.t. the phases
            "value": [float((2 * F).real), float((2 * F).imag)],
            "dvalue_dphase": [[float((2 * d).real), float((2 * d).imag)] for d in (dA, dB, dC)],
        })
        print(f"{name:40s} phases {float(xa):+.6g} {float(xb):+.6g} {float(xc):+.6g}  value {complex(2 * F):.6g}")
    return out


def _I(k):
    k = mp.mpf(k)
    return mp.mpf(1) if k == 0 else (1 - mp.exp(-I1 * k)) / (I1 * k)


def cube_surface():
    v = [[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)]
    idx = {tuple(p): i for i, p in enumerate(v)}
    quads = [  # outward, counter-clockwise seen from outside
        [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)],
        [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
        [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)],
        [(0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)],
        [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)],
        [(1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)],
    ]
    faces = []
    for qd in quads:
        a, b, c, d = (idx[p] for p in qd)
        faces += [[a, b, c], [a, c, d]]
    g = np.random.default_rng(20261203)
    xi = np.concatenate([
        np.zeros((1, 3)),
        np.round(g.normal(0, 12, (96, 3)) * 4) / 4,
        np.array([[3.0, 0, 0], [0, -2.0, 0], [0, 0, 7.5], [1.0, 1.0, 0], [0.25, 0, 0],
                  [40.0, -30.0, 0.5], [-55.25, 12.0, 33.75]]),
        # phase gaps above 1000 rad on the faces spanning x and y
        np.array([[900.25, -700.5, 300.0], [-1200.0, 0.25, 850.75], [640.5, 640.5, -640.5]]),
    ]).astype(np.float32)
    surf, solid = [], []
    for kx, ky, kz in xi.astype(float):
        Ix, Iy, Iz = _I(kx), _I(ky), _I(kz)
        ex, ey, ez = (mp.exp(-I1 * mp.mpf(k)) for k in (kx, ky, kz))
        s = (Ix * Iy * (1 + ez) + Iy * Iz * (1 + ex) + Iz * Ix * (1 + ey)) / 6
        surf.append([float(s.real), float(s.imag)])
        so = Ix * Iy * Iz
        solid.append([float(so.real), float(so.imag)])
    return {"verts": v, "faces": faces, "xi": xi.astype(float).tolist(), "surface": surf,
            "solid": solid, "wide_from": len(xi) - 3}


def main():
    with open(os.path.join(HERE, "single_faces.json"), "w") as fh:
        json.dump(single_faces(), fh, indent=1)
    with open(os.path.join(HERE, "cube_surface.json"), "w") as fh:
        json.dump(cube_surface(), fh, indent=1)


if __name__ == "__main__":
    main()
