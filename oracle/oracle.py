"""Python side of the meshFT oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by the smoke check and by bench.py's CPU-baseline leg;
never by the product package.  Wraps liboracle.so (meshft_oracle.cpp) and
restates the reference's host-side pieces in numpy:

* spatial frequency axes  -- deltamic/render.py:153-172
* Gaussian OTF            -- deltamic/psf.py:15-20
* band compaction         -- deltamic/cuda_class.py:20-23 (Filter > thresh, row-major)
* scatter / render        -- deltamic/cuda_class.py:27-28, deltamic/render.py:177-180
* normalize_tensor        -- deltamic/render.py:5-6
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_l = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64
_int = ctypes.c_int

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_face_accurate.argtypes = [ctypes.c_double] * 3 + [_d]
        L.orc_face_ref_f64.argtypes = [ctypes.c_double] * 3 + [_d]
        L.orc_face_ref_f32.argtypes = [ctypes.c_float] * 3 + [_f]
        L.orc_acc_forward.argtypes = [_d, _i64, _l, _i64, _d, _d, _i64, _d, _int]
        L.orc_acc_backward.argtypes = [_d, _i64, _l, _i64, _d, _d, _i64, _d, _d, _d, _int]
        L.orc_acc_forward_af.argtypes = [_d, _i64, _l, _i64, _d, ctypes.c_double, _d, _i64, _d,
                                         _int]
        L.orc_ref_forward_f32.argtypes = [_f, _i64, _l, _i64, _f, _f, _i64, _f, _int]
        L.orc_ref_forward_f64.argtypes = [_d, _i64, _l, _i64, _d, _d, _i64, _d, _int]
        L.orc_ref_backward_f32.argtypes = [_f, _i64, _l, _i64, _f, _f, _i64, _f, _f, _f, _int]
        L.orc_ref_backward_f64.argtypes = [_d, _i64, _l, _i64, _d, _d, _i64, _d, _d, _d, _int]
        for n in ("orc_acc_forward", "orc_acc_backward", "orc_acc_forward_af",
                  "orc_ref_forward_f32", "orc_ref_forward_f64",
                  "orc_ref_backward_f32", "orc_ref_backward_f64"):
            getattr(L, n).restype = _int
        _lib = L
    return _lib


def _threads(n):
    return int(n) if n else (os.cpu_count() or 1)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a, ct):
    return a.ctypes.data_as(ct)


# --------------------------------------------------------------- single face
def face_accurate(xa, xb, xc):
    """(F, [dF/dxa, dF/dxb, dF/dxc]) for f = -exp(-ix), accurate FP64."""
    out = np.zeros(8)
    lib().orc_face_accurate(float(xa), float(xb), float(xc), _p(out, _d))
    c = out[0::2] + 1j * out[1::2]
    return c[0], c[1:]


def face_ref(xa, xb, xc, dtype=np.float64):
    """Reference-semantics face values (fourier3d_cuda_kernel.cu:74-278)."""
    if dtype == np.float32:
        out = np.zeros(8, np.float32)
        lib().orc_face_ref_f32(float(xa), float(xb), float(xc), _p(out, _f))
    else:
        out = np.zeros(8)
        lib().orc_face_ref_f64(float(xa), float(xb), float(xc), _p(out, _d))
    c = out[0::2].astype(np.float64) + 1j * out[1::2].astype(np.float64)
    return c[0], c[1:]


# --------------------------------------------------------------- mesh level
def _inputs(verts, coeff, xi, fp32_inputs):
    if fp32_inputs:  # the values the HIP path receives, widened exactly
        verts = np.asarray(verts, np.float32)
        coeff = np.asarray(coeff, np.float32)
        xi = np.asarray(xi, np.float32)
    return (_c(verts, np.float64), _c(coeff, np.float64), _c(xi, np.float64).reshape(-1, 3))


def acc_forward(verts, faces, coeff, xi, threads=None, fp32_inputs=True):
    """Accurate ftmesh_list (complex128 [n]).  By default the inputs are first
    rounded to FP32 (what the HIP path sees); fp32_inputs=False keeps FP64
    inputs (finite differences)."""
    V, C, X = _inputs(verts, coeff, xi, fp32_inputs)
    F = _c(faces, np.int64)
    out = np.zeros((len(X), 2))
    rc = lib().orc_acc_forward(_p(V, _d), len(V), _p(F, _l), len(F), _p(C, _d), _p(X, _d), len(X),
                               _p(out, _d), _threads(threads))
    assert rc == 0
    return out[:, 0] + 1j * out[:, 1]


def acc_forward_af(verts, faces, coeff, regint, xi, threads=None, fp32_inputs=True):
    """Accurate autofluorescence-variant ftmesh_list (complex128 [n]); see
    orc_acc_forward_af."""
    V, C, X = _inputs(verts, coeff, xi, fp32_inputs)
    F = _c(faces, np.int64)
    out = np.zeros((len(X), 2))
    rc = lib().orc_acc_forward_af(_p(V, _d), len(V), _p(F, _l), len(F), _p(C, _d),
                                  ctypes.c_double(float(regint)), _p(X, _d), len(X), _p(out, _d),
                                  _threads(threads))
    assert rc == 0
    return out[:, 0] + 1j * out[:, 1]


def acc_backward(verts, faces, coeff, xi, grad, threads=None, return_face=False, fp32_inputs=True):
    """Accurate d/dV sum Re(conj(grad) ftmesh) (float64 [nv, 3])."""
    V, C, X = _inputs(verts, coeff, xi, fp32_inputs)
    F = _c(faces, np.int64)
    g = np.asarray(grad)
    G = _c(np.stack([g.real, g.imag], axis=1), np.float64)
    out = np.zeros((len(V), 3))
    face = np.zeros((len(F), 10)) if return_face else None
    rc = lib().orc_acc_backward(_p(V, _d), len(V), _p(F, _l), len(F), _p(C, _d), _p(X, _d), len(X),
                                _p(G, _d), _p(out, _d), _p(face, _d) if return_face else None,
                                _threads(threads))
    assert rc == 0
    return (out, face) if return_face else out


def ref_forward(verts, faces, coeff, xi, dtype=np.float32, threads=None):
    """Reference-semantics forward (Mesh_NUFT_kernel order and branches)."""
    ct = _f if dtype == np.float32 else _d
    V = _c(verts, dtype)
    F = _c(faces, np.int64)
    C = _c(coeff, dtype)
    X = _c(xi, dtype).reshape(-1, 3)
    out = np.zeros((len(X), 2), dtype)
    fn = lib().orc_ref_forward_f32 if dtype == np.float32 else lib().orc_ref_forward_f64
    assert fn(_p(V, ct), len(V), _p(F, _l), len(F), _p(C, ct), _p(X, ct), len(X), _p(out, ct),
              _threads(threads)) == 0
    return out[:, 0].astype(np.float64) + 1j * out[:, 1].astype(np.float64)


def ref_backward(verts, faces, coeff, xi, grad, ftmesh, dtype=np.float32, threads=None):
    """Reference-semantics adjoint (Grad_Mesh_NUFT_kernel + Additions_kernel)."""
    ct = _f if dtype == np.float32 else _d
    V = _c(verts, dtype)
    F = _c(faces, np.int64)
    C = _c(coeff, dtype)
    X = _c(xi, dtype).reshape(-1, 3)
    g = np.asarray(grad)
    G = _c(np.stack([g.real, g.imag], axis=1), dtype)
    fm = np.asarray(ftmesh)
    FM = _c(np.stack([fm.real, fm.imag], axis=1), dtype)
    out = np.zeros((len(V), 3), dtype)
    fn = lib().orc_ref_backward_f32 if dtype == np.float32 else lib().orc_ref_backward_f64
    assert fn(_p(V, ct), len(V), _p(F, _l), len(F), _p(C, ct), _p(X, ct), len(X), _p(G, ct),
              _p(FM, ct), _p(out, ct), _threads(threads)) == 0
    return out.astype(np.float64)


# --------------------------------------------------------------- host-side restatements
def frequency_axes(box_size, box_shape, dtype=np.float32):
    """xi_d = (2k - (n_d - 1)) * (pi / L_d) in ``dtype`` (render.py:153-172).

    torch computes pi / L as reciprocal(L) * pi in the tensor dtype; this
    does the same so the FP32 axes are bit-identical."""
    axes = []
    for d, n in enumerate(box_shape):
        L = dtype(box_size[d][1]) - dtype(box_size[d][0])
        s = dtype(dtype(1) / L) * dtype(np.pi)
        k = np.arange(n, dtype=dtype)
        axes.append(((dtype(2) * k - dtype(n - 1)) * s).astype(dtype))
    return axes


def gaussian_otf(sigma, xi0, xi1, xi2):
    """exp(-xi^T Sigma xi / 2) on the 'ij' grid (psf.py:15-20), float64."""
    s = np.asarray(sigma, np.float64)
    g0, g1, g2 = np.meshgrid(np.asarray(xi0, np.float64), np.asarray(xi1, np.float64),
                             np.asarray(xi2, np.float64), indexing="ij")
    q = (g0 * (s[0, 0] * g0 + s[0, 1] * g1 + s[0, 2] * g2)
         + g1 * (s[1, 0] * g0 + s[1, 1] * g1 + s[1, 2] * g2)
         + g2 * (s[2, 0] * g0 + s[2, 1] * g1 + s[2, 2] * g2))
    return np.exp(-q / 2)


def band(filter_box, thresh, xi0, xi1, xi2):
    """(linear indices, xi_list) of filter > thresh in row-major order (cuda_class.py:20-23)."""
    f = np.asarray(filter_box)
    mask = f > np.asarray(thresh, dtype=f.dtype)
    idx = np.flatnonzero(mask.reshape(-1))
    n1, n2 = f.shape[1], f.shape[2]
    i0, i1, i2 = idx // (n1 * n2), (idx // n2) % n1, idx % n2
    xi = np.stack([np.asarray(xi0)[i0], np.asarray(xi1)[i1], np.asarray(xi2)[i2]], axis=1)
    return idx, xi


def scatter(vals, idx, shape):
    box = np.zeros(int(np.prod(shape)), dtype=np.complex128)
    box[idx] = vals
    return box.reshape(shape)


def render(ftmesh_box, otf, box_shape):
    """|ifftn(ftmesh * OTF)| (render.py:177-180)."""
    return np.abs(np.fft.ifftn(ftmesh_box * otf, s=list(box_shape)))


def normalize(t, c=0.0):
    return ((t - t.min()) / (t.max() - t.min()) + c) / (1 + c)


def rel_l2(a, b):
    a = np.asarray(a); b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
