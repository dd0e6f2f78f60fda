"""meshFT on the CPU with torch: the ``device='cpu'`` path of ``Fourier3dMesh``.

The reference runs its CPU device through ``MeshFTPy`` (deltamic/render.py:36-109):
a dense torch expression over all [band, faces] pairs, differentiated by
autograd.  This module is the same operator, written for the CPU:

* the same inputs and output: vertices already shifted to the box minimum,
  faces, face coefficients, the three frequency axes, the OTF and the
  threshold; the band is ``|OTF| > thresh`` as in MeshFTPy (render.py:49);
  the result is the full complex box, zero outside the band, in the complex
  type of the axes (complex64 for float32);
* face weights 2 c_t A_t / sum c A with A_t from the cross product (MeshFTPy
  uses Heron's formula, render.py:21-33: the same area);
* each face's value is the second divided difference G_t = f[x_a, x_b, x_c]
  of f(x) = -exp(-ix) at the vertex phases x_j = xi . v_j (the quantity of
  MeshFTPy's five branches, render.py:94-98), evaluated in FP64 in the
  re-based form of the GPU kernels: with the phases sorted lo <= mid <= hi,
  P = lo - mid, Q = hi - mid, D = hi - lo,
      G = exp(-i x_mid) (phi(P) - phi(Q)) / D,
      phi(y) = (exp(-iy) - 1)/y = -sinc(y/2) (sin(y/2) + i cos(y/2)),
  which does not cancel, and its Taylor series in P, Q when D < 1e-2;
* gradients by autograd through that expression, band chunk by band chunk:
  the forward keeps no graph, and the backward re-evaluates each chunk with
  autograd on, so memory stays O(chunk x faces) instead of MeshFTPy's
  O(band x faces) (133 GB at 20,480 faces in a 200^3 box).

Cost is O(band x faces) FP64 arithmetic on the host: meant for the small
meshes the reference's CPU notebook fits (e.g. 1,280 faces in 64^3), not for
the GPU workloads.  It is a device path of its own, chosen only by
``device='cpu'``; GPU tensors never come here.
"""
from __future__ import annotations

import math

import torch

# largest phase gap (rad) handled by the Taylor series of the divided
# difference; the truncation after h_7 is < 1e-2^8 / 10! ~ 3e-23 there
_TAYLOR_D = 1e-2
# FP64 pairs evaluated per chunk (bounds the autograd graph of one chunk)
_CHUNK_PAIRS = 1 << 21


def face_weights(V: torch.Tensor, F: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    """2 c_t A_t / sum_t c_t A_t with A_t = |(b - a) x (c - a)| / 2 (0 if all are 0)."""
    P = V[F]
    n = torch.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], dim=1)
    cA = C * 0.5 * torch.linalg.vector_norm(n, dim=1)
    S = cA.sum()
    return torch.where(S > 0, 2.0 * cA / torch.where(S > 0, S, torch.ones_like(S)),
                       torch.zeros_like(cA))


def _taylor_H(P: torch.Tensor, Q: torch.Tensor) -> torch.Tensor:
    """(phi(P) - phi(Q)) / (Q - P) for small |P|, |Q|: sum_k -(-i)^k h_{k-2}(P, Q)/k!,
    k = 2..9, with the complete homogeneous symmetric polynomials
    h_j = (P + Q) h_{j-1} - P Q h_{j-2}."""
    s, t = P + Q, P * Q
    h = [torch.ones_like(P), s]
    for _ in range(6):
        h.append(s * h[-1] - t * h[-2])
    re = torch.zeros_like(P)
    im = torch.zeros_like(P)
    for k in range(2, 10):
        term = h[k - 2] / math.factorial(k)
        # -(-i)^k: k = 2 -> +1, 3 -> -i, 4 -> -1, 5 -> +i, ...
        r = k % 4
        if r == 0:
            re = re - term
        elif r == 1:
            im = im + term
        elif r == 2:
            re = re + term
        else:
            im = im - term
    return torch.complex(re, im)


def _phi(y: torch.Tensor) -> torch.Tensor:
    """(exp(-iy) - 1)/y = -sinc(y/2) (sin(y/2) + i cos(y/2)), finite at y = 0."""
    k = torch.sinc(y / (2.0 * math.pi))  # sin(y/2)/(y/2)
    h = 0.5 * y
    return torch.complex(-k * torch.sin(h), -k * torch.cos(h))


def face_values(V: torch.Tensor, F: torch.Tensor, xi: torch.Tensor) -> torch.Tensor:
    """G[m, t] = f[x_a, x_b, x_c] of f(x) = -exp(-ix) for frequencies xi [n, 3]
    (FP64) and faces F of vertices V (FP64): complex128 [n, nf]."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    xa = xi @ a.T                      # vertex a's phase
    p = xi @ (b - a).T                 # phase differences from edge vectors
    q = xi @ (c - a).T
    srt = torch.sort(torch.stack((torch.zeros_like(p), p, q)), dim=0).values
    lo, mid, hi = srt[0], srt[1], srt[2]
    P, Q, D = lo - mid, hi - mid, hi - lo
    small = D < _TAYLOR_D
    Dsafe = torch.where(small, torch.ones_like(D), D)
    H = torch.where(small, _taylor_H(P, Q), (_phi(P) - _phi(Q)) / Dsafe)
    x = xa + mid
    return torch.complex(torch.cos(x), -torch.sin(x)) * H


def _chunks(n: int, nf: int):
    step = max(1, _CHUNK_PAIRS // max(nf, 1))
    return [(s, min(n, s + step)) for s in range(0, n, step)]


def _chunk_values(V, F, C, xi, regint):
    """ftmesh on the frequencies xi (complex128 [n]).  regint is None for the
    surface transform sum_t w_t G_t, w_t = 2 c_t A_t / sum c A; else the
    autofluorescence transform (render.py:228-318)
        sum_t G_t (w_t + i (xi . nor2_t) / |xi|^2),  w_t = 2 A_t / sum A,
        nor2_t = (b - a) x (c - a) regint / regvol,  regvol = -sum_t a . (b x c) / 6,
    whose interior term at xi = 0 is its limit -regint (the reference divides
    by zero there)."""
    G = face_values(V, F, xi)
    w = face_weights(V, F, C).to(torch.complex128)
    out = G @ w
    if regint is None:
        return out
    P = V[F]
    nor = torch.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], dim=1)
    regvol = -(P[:, 0] * torch.cross(P[:, 1], P[:, 2], dim=1)).sum() / 6.0
    ok = regvol != 0
    scale = torch.where(ok, regint / torch.where(ok, regvol, torch.ones_like(regvol)),
                        torch.zeros_like(regvol))
    s = (G * (xi @ (nor * scale).T)).sum(1)  # sum_t G_t (xi . nor2_t)
    q = (xi * xi).sum(1)
    zero = q == 0
    inv = torch.where(zero, torch.zeros_like(q), 1.0 / torch.where(zero, torch.ones_like(q), q))
    lim = torch.where(zero & ok, torch.full_like(q, -float(regint)), torch.zeros_like(q))
    return out + torch.complex(lim, torch.zeros_like(lim)) + 1j * inv * s


def band_values(V: torch.Tensor, F: torch.Tensor, C: torch.Tensor, xi: torch.Tensor,
                regint=None) -> torch.Tensor:
    """ftmesh on the band frequencies xi [n, 3] (complex128 [n]), chunk by chunk."""
    out = torch.empty(xi.shape[0], dtype=torch.complex128)
    for s, e in _chunks(xi.shape[0], F.shape[0]):
        out[s:e] = _chunk_values(V, F, C, xi[s:e], regint)
    return out


class MeshFTCPU(torch.autograd.Function):
    """ftmesh on the band (complex128 [n]) from FP64 vertices, differentiable
    w.r.t. the vertices; the backward re-evaluates one band chunk at a time
    under autograd (no [band, faces] graph is kept).  regint: None for the
    surface transform, a number for the autofluorescence variant."""

    @staticmethod
    def forward(ctx, V, F, C, xi, regint=None):
        ctx.save_for_backward(V, F, C, xi)
        ctx.regint = regint
        with torch.no_grad():
            return band_values(V, F, C, xi, regint)

    @staticmethod
    def backward(ctx, g):
        V, F, C, xi = ctx.saved_tensors
        grad = torch.zeros_like(V)
        g = g.to(torch.complex128)
        for s, e in _chunks(xi.shape[0], F.shape[0]):
            with torch.enable_grad():
                Vl = V.detach().requires_grad_(True)
                val = _chunk_values(Vl, F, C, xi[s:e], ctx.regint)
                (gv,) = torch.autograd.grad(val, Vl, g[s:e])
            grad += gv
        return grad, None, None, None, None


def meshft_cpu(Verts, Faces, Faces_coeff, xi0, xi1, xi2, OTF, narrowband_thresh):
    """The ``device='cpu'`` forward of ``Fourier3dMesh`` (render.py:142-144):
    complex box of the axes' complex type, zero outside ``|OTF| > thresh``."""
    mask = torch.abs(OTF) > narrowband_thresh
    g0, g1, g2 = torch.meshgrid(xi0, xi1, xi2, indexing="ij")
    xi = torch.stack((g0[mask], g1[mask], g2[mask]), dim=1).to(torch.float64)
    vals = MeshFTCPU.apply(Verts.to(torch.float64), Faces.long(),
                           Faces_coeff.to(torch.float64), xi)
    ctype = torch.complex128 if xi0.dtype == torch.float64 else torch.complex64
    out = torch.zeros(tuple(mask.shape), dtype=ctype)
    out = out.index_put((mask,), vals.to(ctype))
    return out


def meshft_af_cpu(OTF, narrowband_thresh, regint, xi0, xi1, xi2, Verts, Faces):
    """``fourier_from_mesh_narrowband_autofluorescence`` on host tensors
    (render.py:228-318): band ``OTF > thresh`` (this function of the reference
    takes no abs), phases of the given vertices, complex box of the axes'
    complex type, zero outside the band."""
    mask = OTF > narrowband_thresh
    g0, g1, g2 = torch.meshgrid(xi0, xi1, xi2, indexing="ij")
    xi = torch.stack((g0[mask], g1[mask], g2[mask]), dim=1).to(torch.float64)
    F = Faces.long()
    vals = MeshFTCPU.apply(Verts.to(torch.float64), F,
                           torch.ones(F.shape[0], dtype=torch.float64), xi, float(regint))
    ctype = torch.complex128 if xi0.dtype == torch.float64 else torch.complex64
    out = torch.zeros(tuple(mask.shape), dtype=ctype)
    return out.index_put((mask,), vals.to(ctype))
