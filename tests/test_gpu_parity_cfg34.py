"""GPU parity on the fit workload (config 4) and, pair by pair, on the
precision-stress workload (config 3), against the CPU oracle.

Tolerances: forward rel-L2 <= 1e-5, gradients rel-L2 <= 1e-4.

Config 4 (icosphere s4 with SH l <= 8, 128^3 box, band of 1.17M
frequencies): the HIP forward and adjoint against the FP64 oracle on 4096
seeded band frequencies, and the HIP adjoint against central differences of
the FP64 oracle forward at h = 1e-3 on 64 seeded vertices (seed 4000 + 2)
over 512 seeded band frequencies.  At this band's |xi| <= 152 the two-point
central difference carries a truncation error of (h |xi|)^2 / 6 ~ 4e-3
relative; the five-point central difference at the same h (truncation
(h |xi|)^4 / 30 ~ 2e-5) is the one checked against 1e-4, the two-point one
against its own truncation bound.

Config 3 (410 needle faces, anisotropic OTF, 256 x 256 x 128): one-face meshes
(face weight 2, so the value is 2 F_t) give every (face, frequency) pair's
value and gradient separately.  Pairs are classified from the FP32 inputs:
  ref-degenerate  some phase difference |x_j - x_k| <= sqrt(eps_f32), where the
                  reference switches to its small-difference Taylor branches
                  (fourier3d_cuda_kernel.cu:250-271);
  taylor          largest phase gap D < 0.1 rad (the HIP kernels' Taylor branch);
  clamped         |P| or |Q| < 1.3e-8 rad (the HIP kernels clamp them);
  generic         everything else;
and the error is reported per class (rel-L2 over the class's pairs).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import threads
from deltamic_amd import synthetic as syn

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
BWD_TOL = 1e-4


def _band(w, device):
    from deltamic_amd import generate_gaussian_psf, spatial_frequency_axes
    from deltamic_amd.hip_class import compute_band
    bs = torch.tensor(w.box_size, dtype=torch.float32, device=device)
    axes = spatial_frequency_axes(bs, w.box_shape, torch.float32, device)
    otf = generate_gaussian_psf(torch.tensor(w.sigma, device=device), *axes)
    return bs, compute_band(otf, w.thresh, *axes)


def _subset(w, band, count, device, seed):
    sel = np.sort(syn.rng(seed).choice(band.idx.numel(), count, replace=False))
    return band.xi_list[torch.tensor(sel, device=device)].contiguous()


def _grad(n, seed):
    g = syn.rng(seed)
    return (g.standard_normal(n) + 1j * g.standard_normal(n)).astype(np.complex64)


def _hip(Vs, faces, coeff, xi, g, device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    V = torch.tensor(np.ascontiguousarray(Vs, np.float32), device=device)
    Fc = torch.tensor(np.ascontiguousarray(faces, np.int64), device=device)
    C = torch.tensor(np.ascontiguousarray(coeff, np.float32), device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, xi)
    gv = None
    if g is not None:
        gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(g, device=device), ft, xi)
        gv = gv.cpu().numpy().astype(np.float64)
    return ft.cpu().numpy().astype(np.complex128), gv


def test_config4_forward_and_adjoint_vs_oracle(device):
    w = syn.config4(0)
    bs, band = _band(w, device)
    assert 1.1e6 < band.idx.numel() < 1.25e6
    xi = _subset(w, band, 4096, device, w.seed + 1)
    Vs = (torch.tensor(w.verts) - bs.cpu()[:, 0]).numpy()
    g = _grad(4096, w.seed + 4)
    ft, gv = _hip(Vs, w.faces, w.coeff, xi, g, device)
    xi_np = xi.cpu().numpy()
    ef = O.rel_l2(ft, O.acc_forward(Vs, w.faces, w.coeff, xi_np, threads=threads()))
    eb = O.rel_l2(gv, O.acc_backward(Vs, w.faces, w.coeff, xi_np, g, threads=threads()))
    print(f"config 4 (5120 faces, 4096 of {band.idx.numel()} band freqs): forward rel-L2 "
          f"{ef:.3e}, adjoint rel-L2 {eb:.3e}")
    assert ef <= FWD_TOL and eb <= BWD_TOL


def _unnormalised(V, faces, coeff, xi):
    """(sum_t c_t A_t F_t(xi) [n] complex, sum_t c_t A_t) with the FP64 oracle:
    its forward is sum_t 2 c_t A_t F_t / sum_s c_s A_s."""
    P = V[faces]
    A = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    S = float(np.dot(coeff, A))
    ft = O.acc_forward(V, faces, coeff, xi, threads=threads(), fp32_inputs=False)
    return ft * (S / 2.0), S


def test_config4_adjoint_vs_oracle_central_differences(device):
    """L(V) = sum_m Re(conj(g_m) ftmesh_m(V)); dL/dv from the HIP adjoint
    against central differences of the FP64 oracle forward.  A vertex moves
    only its incident faces and the area sum S, so
    ftmesh(V') = 2 (U(V) - U_inc(V) + U_inc(V')) / S(V'), U = sum_t c_t A_t F_t,
    with U_inc over the faces incident to the moved vertex."""
    w = syn.config4(0)
    bs, band = _band(w, device)
    xi = _subset(w, band, 512, device, w.seed + 1)
    xi_np = xi.cpu().numpy().astype(np.float64)
    V32 = (torch.tensor(w.verts) - bs.cpu()[:, 0]).numpy()
    V = V32.astype(np.float64)
    faces, coeff = w.faces, w.coeff.astype(np.float64)
    g = _grad(512, w.seed + 4)
    _, gv = _hip(V32, faces, w.coeff, xi, g, device)
    U, S = _unnormalised(V, faces, coeff, xi_np)
    verts = np.sort(syn.rng(w.seed + 2).choice(len(V), 64, replace=False))
    h = 1e-3
    fd2 = np.zeros((64, 3))
    fd4 = np.zeros((64, 3))
    for i, v in enumerate(verts):
        inc = np.flatnonzero((faces == v).any(axis=1))
        U0, S0 = _unnormalised(V, faces[inc], coeff[inc], xi_np)
        for d in range(3):
            Ls = {}
            for s in (-2, -1, 1, 2):
                Vp = V.copy()
                Vp[v, d] += s * h
                Ui, Si = _unnormalised(Vp, faces[inc], coeff[inc], xi_np)
                ft = 2.0 * (U - U0 + Ui) / (S - S0 + Si)
                Ls[s] = float(np.sum(g.real * ft.real + g.imag * ft.imag))
            fd2[i, d] = (Ls[1] - Ls[-1]) / (2 * h)
            fd4[i, d] = (-Ls[2] + 8 * Ls[1] - 8 * Ls[-1] + Ls[-2]) / (12 * h)
    an = gv[verts]
    e4 = np.linalg.norm(an - fd4) / np.linalg.norm(fd4)
    e2 = np.linalg.norm(an - fd2) / np.linalg.norm(fd2)
    kmax = float(np.linalg.norm(xi_np, axis=1).max())
    print(f"config 4 adjoint vs oracle central differences (64 vertices, h = {h}, 512 freqs, "
          f"|xi| <= {kmax:.0f}): five-point rel-L2 {e4:.3e}, two-point rel-L2 {e2:.3e} "
          f"(two-point truncation bound {(h * kmax) ** 2 / 6:.1e})")
    assert e4 <= BWD_TOL
    assert e2 <= (h * kmax) ** 2 / 6


def _classes(Vs, faces, xi):
    """Per (face, frequency) class labels from the FP32 inputs (see module doc)."""
    P = Vs[faces].astype(np.float64)            # [nf, 3, 3]
    X = xi.astype(np.float64)                   # [n, 3]
    ph = np.einsum("tjd,md->tmj", P, X)         # phases [nf, n, 3]
    p = ph[:, :, 1] - ph[:, :, 0]
    q = ph[:, :, 2] - ph[:, :, 0]
    lo = np.minimum(0.0, np.minimum(p, q))
    hi = np.maximum(0.0, np.maximum(p, q))
    mid = np.median(np.stack([np.zeros_like(p), p, q]), axis=0)
    Pm, Qm, D = lo - mid, hi - mid, hi - lo
    dmin = np.minimum(np.minimum(np.abs(p), np.abs(q)), np.abs(q - p))
    cls = np.full(p.shape, "generic", dtype=object)
    cls[np.minimum(-Pm, Qm) < 1.3e-8] = "clamped"
    cls[D < 0.1] = "taylor"
    ref_deg = dmin <= np.sqrt(np.finfo(np.float32).eps)
    return cls, ref_deg


def test_config3_error_by_branch(device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config3()
    bs, band = _band(w, device)
    xi = _subset(w, band, 4096, device, w.seed + 1)
    xi_np = xi.cpu().numpy()
    Vs = (torch.tensor(w.verts) - bs.cpu()[:, 0]).numpy()
    needles = w.extra["needle_faces"]
    others = np.setdiff1d(np.arange(len(w.faces)), needles)
    faces = np.concatenate([needles, np.sort(syn.rng(w.seed + 5).choice(others, len(needles),
                                                                         replace=False))])
    cls, ref_deg = _classes(Vs, w.faces[faces], xi_np)
    g = _grad(4096, w.seed + 4)
    names = ("ref-degenerate", "taylor", "clamped", "generic")
    acc = {k: np.zeros(4) for k in names}      # fwd err^2, fwd ref^2, grad err^2, grad ref^2
    count = {k: 0 for k in names}
    one = np.array([[0, 1, 2]], np.int64)
    c1 = np.ones(1, np.float32)
    Fc = torch.tensor(one, device=device)
    C = torch.tensor(c1, device=device)
    for i, t in enumerate(faces):
        v3 = np.ascontiguousarray(Vs[w.faces[t]])
        V = torch.tensor(v3, device=device)
        ft = forward_fourier3d_cpp(V, Fc, C, xi).cpu().numpy().astype(np.complex128)
        ref = O.acc_forward(v3, one, c1, xi_np, threads=threads())
        masks = {"ref-degenerate": ref_deg[i], "taylor": cls[i] == "taylor",
                 "clamped": cls[i] == "clamped",
                 "generic": (cls[i] == "generic") & ~ref_deg[i]}
        for k, m in masks.items():
            if not m.any():
                continue
            count[k] += int(m.sum())
            acc[k][0] += np.sum(np.abs(ft[m] - ref[m]) ** 2)
            acc[k][1] += np.sum(np.abs(ref[m]) ** 2)
            gm = np.where(m, g, 0).astype(np.complex64)
            gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(gm, device=device),
                                        torch.tensor(ft.astype(np.complex64), device=device), xi)
            refg = O.acc_backward(v3, one, c1, xi_np, gm, threads=threads())
            acc[k][2] += np.sum((gv.cpu().numpy() - refg) ** 2)
            acc[k][3] += np.sum(refg ** 2)
    for k in names:
        if count[k] == 0:
            print(f"config 3 {k}: no pairs")
            continue
        a = acc[k]
        ef, eg = np.sqrt(a[0] / a[1]), np.sqrt(a[2] / max(a[3], 1e-300))
        print(f"config 3 {k}: {count[k]} pairs, forward rel-L2 {ef:.3e}, gradient rel-L2 {eg:.3e}")
        assert ef <= FWD_TOL and eg <= BWD_TOL
    # the needle faces put many pairs in the reference's small-difference branches
    assert count["ref-degenerate"] > 10000 and count["taylor"] > 1000
