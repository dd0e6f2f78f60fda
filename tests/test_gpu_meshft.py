"""GPU parity of the HIP meshFT path against the CPU oracle (oracle/).

Tolerances (rel-L2 over the band or over all vertices):
  forward  <= 1e-5 vs the accurate FP64 oracle,
  adjoint  <= 1e-4 vs the accurate FP64 oracle,
  band / scatter / gather / face-index bookkeeping: exact.
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


def _mesh(w, device):
    V = torch.tensor(w.verts, device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    return V, Fc, C


def _band(w, device):
    """GPU band of the workload's Gaussian OTF, checked against numpy on the same OTF."""
    from deltamic_amd import generate_gaussian_psf, spatial_frequency_axes
    from deltamic_amd.hip_class import compute_band
    bs = torch.tensor(w.box_size, dtype=torch.float32, device=device)
    axes = spatial_frequency_axes(bs, w.box_shape, torch.float32, device)
    otf = generate_gaussian_psf(torch.tensor(w.sigma, device=device), *axes)
    band = compute_band(otf, w.thresh, *axes)
    return bs, axes, otf, band


def _shifted(w, bs):
    # Verts - box_size[:, 0] in FP32, as Fourier3dMesh.forward does (render.py:148)
    return (torch.tensor(w.verts) - bs.cpu()[:, 0]).numpy()


def test_band_and_otf_match_numpy(device):
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    ax_np = O.frequency_axes(bs.cpu().numpy(), w.box_shape)
    for a, b in zip(axes, ax_np):
        assert np.array_equal(a.cpu().numpy(), b)
    ref = O.gaussian_otf(w.sigma.astype(np.float64), *ax_np)
    got = otf.cpu().numpy().astype(np.float64)
    # FP32 exponent q has relative error ~1e-7, so exp(-q/2) has ~1e-7 (1 + q) relative error
    live = ref > 1e-30
    q = -2.0 * np.log(ref[live])
    relerr = np.abs(got[live] - ref[live]) / ref[live]
    assert np.all(relerr <= 2e-7 * (4.0 + q)), relerr.max()
    assert np.all(got[~live] < 1e-29)
    idx, xi = O.band(otf.cpu().numpy(), np.float32(w.thresh), *ax_np)
    assert np.array_equal(band.idx.cpu().numpy().astype(np.int64), idx)
    assert np.array_equal(band.xi_list.cpu().numpy(), xi)
    assert 20000 < len(idx) < 24000


def test_band_random_filter_exact(device):
    from deltamic_amd.hip_class import compute_band
    g = np.random.default_rng(7)
    for shape in [(1, 1, 1), (3, 5, 7), (33, 17, 65), (64, 64, 64)]:
        f = g.random(shape, dtype=np.float32)
        axes = [torch.tensor(g.standard_normal(n).astype(np.float32), device=device) for n in shape]
        band = compute_band(torch.tensor(f, device=device), 0.5, *axes)
        idx, xi = O.band(f, np.float32(0.5), *[a.cpu().numpy() for a in axes])
        assert np.array_equal(band.idx.cpu().numpy().astype(np.int64), idx)
        assert np.array_equal(band.xi_list.cpu().numpy(), xi)


def test_forward_cfg1_full_band(device):
    from deltamic_amd.fourier3d_cpp import forward_fourier3d_cpp
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    Vs = _shifted(w, bs)
    out = forward_fourier3d_cpp(torch.tensor(Vs, device=device), torch.tensor(w.faces, device=device),
                                torch.tensor(w.coeff, device=device), band.xi_list)
    xi = band.xi_list.cpu().numpy()
    ref = O.acc_forward(Vs, w.faces, w.coeff, xi, threads=threads())
    got = out.cpu().numpy().astype(np.complex128)
    err = O.rel_l2(got, ref)
    # least-squares complex scale between the two: the rounding errors of
    # 22,880 frequencies average out in it, a systematic error (a constant of
    # the kernels off by a relative 1e-7, say) does not
    scale = np.vdot(ref, got) / np.vdot(ref, ref)
    print(f"cfg1 forward rel-L2 vs FP64 oracle: {err:.3e} (n={len(xi)}); "
          f"least-squares scale - 1: {abs(scale - 1):.2e}")
    assert err <= FWD_TOL
    assert abs(scale - 1) <= 1e-7


def test_forward_cfg2_subset(device):
    """20480 faces, 200^3 box: 4096 band frequencies drawn by seed + 1."""
    from deltamic_amd.fourier3d_cpp import forward_fourier3d_cpp
    w = syn.config2()
    bs, axes, otf, band = _band(w, device)
    n = band.idx.numel()
    # (4/3) pi (r_band / (2 pi / L))^3 with r_band = sqrt(2 ln(1e5) / 1e-3) and
    # this shape's box edge L ~ 2.64: ~1.08e6 band frequencies
    assert 1.0e6 < n < 1.2e6
    sel = np.sort(syn.rng(w.seed + 1).choice(n, 4096, replace=False))
    xi_sub = band.xi_list[torch.tensor(sel, device=device)].contiguous()
    Vs = _shifted(w, bs)
    out = forward_fourier3d_cpp(torch.tensor(Vs, device=device), torch.tensor(w.faces, device=device),
                                torch.tensor(w.coeff, device=device), xi_sub)
    ref = O.acc_forward(Vs, w.faces, w.coeff, xi_sub.cpu().numpy(), threads=threads())
    err = O.rel_l2(out.cpu().numpy(), ref)
    print(f"cfg2 forward rel-L2 vs FP64 oracle (4096 freqs): {err:.3e}")
    assert err <= FWD_TOL


def _grad(n, seed):
    g = syn.rng(seed)
    return (g.standard_normal(n) + 1j * g.standard_normal(n)).astype(np.complex64)


def test_backward_cfg1_full_band(device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    Vs = _shifted(w, bs)
    V = torch.tensor(Vs, device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    n = band.idx.numel()
    g = _grad(n, w.seed + 4)
    ft = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
    gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(g, device=device), ft, band.xi_list)
    ref = O.acc_backward(Vs, w.faces, w.coeff, band.xi_list.cpu().numpy(), g, threads=threads())
    got = gv.cpu().numpy().astype(np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    err = O.rel_l2(got, ref)
    scale = np.dot(ref, got) / np.dot(ref, ref)
    print(f"cfg1 adjoint rel-L2 vs FP64 oracle: {err:.3e}; least-squares scale - 1: "
          f"{abs(scale - 1):.2e}")
    assert err <= BWD_TOL
    assert abs(scale - 1) <= 1e-6


def test_backward_cfg2_subset(device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config2()
    bs, axes, otf, band = _band(w, device)
    n = band.idx.numel()
    sel = np.sort(syn.rng(w.seed + 1).choice(n, 4096, replace=False))
    xi_sub = band.xi_list[torch.tensor(sel, device=device)].contiguous()
    Vs = _shifted(w, bs)
    V = torch.tensor(Vs, device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    g = _grad(4096, w.seed + 4)
    ft = forward_fourier3d_cpp(V, Fc, C, xi_sub)
    gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(g, device=device), ft, xi_sub)
    ref = O.acc_backward(Vs, w.faces, w.coeff, xi_sub.cpu().numpy(), g, threads=threads())
    err = O.rel_l2(gv.cpu().numpy(), ref)
    print(f"cfg2 adjoint rel-L2 vs FP64 oracle (4096 freqs): {err:.3e}")
    assert err <= BWD_TOL


def test_wide_phase_faces(device):
    """Faces whose phase gap exceeds 16 rad take the range-reduced variant;
    here gaps reach thousands of rad, beyond the hardware sin/cos domain of
    256 revolutions.
    Vertices are multiples of 1/8 and frequencies multiples of 1/4, so every
    FP32 edge vector and phase difference is exact and the usual tolerances
    apply.  Large random triangles (phase gaps of thousands of rad) are mixed
    with small ones (tens of rad) in the same waves, which then take the
    reduced arguments for all their lanes."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    g = syn.rng(77)
    big = (g.integers(-48, 49, size=(120, 3)) / 8.0).astype(np.float32)
    near = [(big[:32] + g.integers(-2, 3, size=(32, 3)) / 8.0).astype(np.float32) for _ in range(2)]
    Vs = np.concatenate([big] + near)
    faces = [g.choice(120, 3, replace=False) for _ in range(96)]
    faces += [(i, 120 + i, 152 + i) for i in range(32)]
    faces = np.array(faces, dtype=np.int64)
    e1 = Vs[faces[:, 1]] - Vs[faces[:, 0]]
    e2 = Vs[faces[:, 2]] - Vs[faces[:, 0]]
    keep = np.linalg.norm(np.cross(e1, e2), axis=1) > 1e-3  # drop collinear draws
    faces, e1, e2, nbig = faces[keep], e1[keep], e2[keep], int(keep[:96].sum())
    coeff = g.uniform(0.5, 1.5, len(faces)).astype(np.float32)
    xi = (g.integers(-1200, 1201, size=(768, 3)) / 4.0).astype(np.float32)
    # phase gaps per pair: most large faces exceed 1000 rad, the small ones stay below
    p, q = xi.astype(np.float64) @ e1.T, xi.astype(np.float64) @ e2.T
    gap = np.maximum.reduce([np.abs(p), np.abs(q), np.abs(p - q)])
    assert (gap[:, :nbig] > 1000).mean() > 0.5 and (gap[:, nbig:] < 1000).all()
    V = torch.tensor(Vs, device=device)
    Fc = torch.tensor(faces, device=device)
    C = torch.tensor(coeff, device=device)
    X = torch.tensor(xi, device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, X)
    ref = O.acc_forward(Vs, faces, coeff, xi, threads=threads())
    ferr = O.rel_l2(ft.cpu().numpy(), ref)
    gr = _grad(len(xi), 78)
    gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(gr, device=device), ft, X)
    refb = O.acc_backward(Vs, faces, coeff, xi, gr, threads=threads())
    berr = O.rel_l2(gv.cpu().numpy(), refb)
    print(f"wide-phase faces: forward rel-L2 {ferr:.3e}, adjoint rel-L2 {berr:.3e}")
    assert ferr <= FWD_TOL
    assert berr <= BWD_TOL


def test_multimaterial_doublet(device):
    """Two-cell multi-material mesh (non-manifold interface loop, valence up to
    9) with Faces_coeff from its labels (1 exterior, 2 cell-cell), the
    reference's multi-material example: forward and adjoint on a 48^3 band
    against the oracle."""
    from deltamic_amd import compute_box_size
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    from deltamic_amd.hip_class import compute_band
    from deltamic_amd import generate_gaussian_psf, spatial_frequency_axes
    from deltamic_amd.meshio import multimaterial_coefficients
    v, fmm = syn.doublet()
    faces = np.ascontiguousarray(fmm[:, :3])
    coeff = multimaterial_coefficients(fmm)
    box = compute_box_size(v, offset=0.2)
    bs = torch.tensor(box, dtype=torch.float32, device=device)
    axes = spatial_frequency_axes(bs, (48, 48, 48), torch.float32, device)
    otf = generate_gaussian_psf(torch.tensor(2e-3 * np.eye(3), dtype=torch.float32, device=device),
                                *axes)
    band = compute_band(otf, 1e-5, *axes)
    Vs = (torch.tensor(v, dtype=torch.float32) - bs.cpu()[:, 0]).numpy()
    V = torch.tensor(Vs, device=device)
    Fc = torch.tensor(faces, device=device)
    C = torch.tensor(coeff, device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
    xi = band.xi_list.cpu().numpy()
    ferr = O.rel_l2(ft.cpu().numpy(), O.acc_forward(Vs, faces, coeff, xi, threads=threads()))
    g = _grad(len(xi), 6004)
    gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(g, device=device), ft, band.xi_list)
    berr = O.rel_l2(gv.cpu().numpy(), O.acc_backward(Vs, faces, coeff, xi, g, threads=threads()))
    print(f"multi-material doublet (n={len(xi)}): forward rel-L2 {ferr:.3e}, adjoint {berr:.3e}")
    assert ferr <= FWD_TOL
    assert berr <= BWD_TOL


def test_bitwise_deterministic(device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    V = torch.tensor(_shifted(w, bs), device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    g = torch.tensor(_grad(band.idx.numel(), 5), device=device)
    a = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
    b = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
    assert torch.equal(a.view(torch.float32), b.view(torch.float32))
    ga = backward_fourier3d_cpp(V, Fc, C, g, a, band.xi_list)
    gb = backward_fourier3d_cpp(V, Fc, C, g, a, band.xi_list)
    assert torch.equal(ga, gb)


def test_precision_stress_cfg3_needles(device):
    """Config 3: 2% near-degenerate faces (one edge 1e-4 of its length),
    anisotropic OTF, threshold 1e-7, 256 x 256 x 128.  Errors on 4096 seeded
    band frequencies, for the whole mesh and separately for the needle faces
    alone (Faces_coeff = 1 on needles, 0 elsewhere)."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config3()
    bs, axes, otf, band = _band(w, device)
    n = band.idx.numel()
    assert 6.0e5 < n < 1.2e6
    sel = np.sort(syn.rng(w.seed + 1).choice(n, 4096, replace=False))
    xi_sub = band.xi_list[torch.tensor(sel, device=device)].contiguous()
    xi_np = xi_sub.cpu().numpy()
    Vs = _shifted(w, bs)
    V = torch.tensor(Vs, device=device)
    Fc = torch.tensor(w.faces, device=device)
    g = _grad(4096, w.seed + 4)
    gt = torch.tensor(g, device=device)
    results = {}
    needle_coeff = np.zeros(len(w.faces), np.float32)
    needle_coeff[w.extra["needle_faces"]] = 1.0
    for label, coeff in (("all faces", w.coeff), ("needle faces only", needle_coeff)):
        C = torch.tensor(coeff, device=device)
        ft = forward_fourier3d_cpp(V, Fc, C, xi_sub)
        gv = backward_fourier3d_cpp(V, Fc, C, gt, ft, xi_sub)
        ref = O.acc_forward(Vs, w.faces, coeff, xi_np, threads=threads())
        refg = O.acc_backward(Vs, w.faces, coeff, xi_np, g, threads=threads())
        results[label] = (O.rel_l2(ft.cpu().numpy(), ref), O.rel_l2(gv.cpu().numpy(), refg))
        print(f"cfg3 {label}: forward rel-L2 {results[label][0]:.3e}, adjoint rel-L2 "
              f"{results[label][1]:.3e}")
    for fe, ge in results.values():
        assert fe <= FWD_TOL
        assert ge <= BWD_TOL


def test_hermitian_halving_matches_full_band(device):
    """Half-band evaluation + mirror equals the full-band evaluation."""
    from deltamic_amd import hip_class
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    assert band.symmetric
    xi = band.xi_list.cpu().numpy()
    n = len(xi)
    assert np.array_equal(xi[::-1], -xi)  # entry n-1-j is exactly -xi_j
    V = torch.tensor(_shifted(w, bs), device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    full = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
    half = hip_class.band_forward(V, Fc, C, band)
    h = half.cpu().numpy()
    assert np.array_equal(h[::-1][: n // 2], np.conj(h[: n // 2]))
    assert O.rel_l2(h, full.cpu().numpy()) <= 1e-6
    g = torch.tensor(_grad(n, 77), device=device)
    gfull = backward_fourier3d_cpp(V, Fc, C, g, full, band.xi_list)
    ghalf = hip_class.band_backward(V, Fc, C, g, half, band)
    assert O.rel_l2(ghalf.cpu().numpy(), gfull.cpu().numpy()) <= 1e-6


def test_asymmetric_band_is_not_halved(device):
    from deltamic_amd.hip_class import compute_band
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    skew = otf * (1.0 + 0.9 * torch.tanh(axes[0] / 50.0))[:, None, None]
    b2 = compute_band(skew.contiguous(), w.thresh, *axes)
    assert not b2.symmetric


def test_closure_halving_on_fftshift_otf(device):
    """The reference fit loop's OTF, |fftshift(fftn(PSF))| of a learnable
    Gaussian PSF (Gradient_Descent.ipynb cell 2, threshold 1e-2), is symmetric
    under k <-> n-k, not under the grid mirror k <-> n-1-k, so its band is not
    centrosymmetric.  Forward and adjoint through the half-evaluated symmetric
    closure equal the full-band evaluation (rel-L2 <= 1e-6)."""
    from deltamic_amd import hip_class
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    from deltamic_amd.psf import Gaussian_psf
    w = syn.config1()
    bs, axes, _, _ = _band(w, device)
    psf = Gaussian_psf(w.box_shape, torch.tensor(w.box_size)[:, 1], sigma=1e3, device=device)
    with torch.no_grad():
        otf = torch.abs(torch.fft.fftshift(torch.fft.fftn(psf.forward()))).contiguous()
    band = hip_class.compute_band(otf, 1e-2, *axes)
    assert not band.symmetric and band.closure is not None and band.closure.symmetric
    cb = band.closure
    print(f"fftshift OTF band {band.n}, closure {cb.n} ({cb.n / band.n:.3f}x), evaluated "
          f"{cb.half}")
    assert band.n < cb.n < 1.3 * band.n
    # the closure holds every band entry at pos, with the same frequency
    assert torch.equal(cb.idx[band.pos.long()], band.idx)
    assert torch.equal(cb.xi_list[band.pos.long()], band.xi_list)
    V = torch.tensor(_shifted(w, bs), device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    full = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
    half = hip_class.band_forward(V, Fc, C, band)
    ef = O.rel_l2(half.cpu().numpy(), full.cpu().numpy())
    g = torch.tensor(_grad(band.n, 78), device=device)
    gfull = backward_fourier3d_cpp(V, Fc, C, g, full, band.xi_list)
    ghalf = hip_class.band_backward(V, Fc, C, g, half, band)
    eb = O.rel_l2(ghalf.cpu().numpy(), gfull.cpu().numpy())
    print(f"closure vs full band: forward rel-L2 {ef:.2e}, adjoint rel-L2 {eb:.2e}")
    assert ef <= 1e-6 and eb <= 1e-6


def test_float64_inputs_config1(device):
    """float64 vertices and frequencies (the reference's double dispatch,
    fourier3d_cuda_kernel.cu:447,508): complex128 / float64 results within
    the float32 path's tolerances of the FP64 oracle on the same FP64 inputs,
    through the native module and through Fourier3dMesh(dtype=float64)."""
    from deltamic_amd import Fourier3dMesh, generate_gaussian_psf
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    box = torch.tensor(w.box_size)
    mesh_ft = Fourier3dMesh(box, w.box_shape, device=device, dtype=torch.float64)
    otf = generate_gaussian_psf(torch.tensor(w.sigma, device=device, dtype=torch.float64),
                                mesh_ft.xi0, mesh_ft.xi1, mesh_ft.xi2)
    assert otf.dtype == torch.float64
    mesh_ft.OTF = otf
    mesh_ft.narrowband_thresh = w.thresh
    # FP64 vertices off the FP32 grid: jitter at 1e-9 below FP32 resolution
    V64 = (torch.tensor(w.verts, dtype=torch.float64)
           + 1e-9 * torch.tensor(syn.rng(9).normal(size=w.verts.shape)))
    Vd = V64.to(device).requires_grad_(True)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.ones(len(w.faces), dtype=torch.float64, device=device)
    ft = mesh_ft(Vd, Fc, C)
    assert ft.dtype == torch.complex128
    from deltamic_amd.hip_class import band_for
    band = band_for(otf, w.thresh, mesh_ft.xi0, mesh_ft.xi1, mesh_ft.xi2)
    xi64 = band.freqs(torch.float64).cpu().numpy()
    Vs = (V64 - mesh_ft.box_size.cpu()[:, 0]).numpy()
    ref = O.acc_forward(Vs, w.faces, np.ones(len(w.faces)), xi64, threads=threads(),
                        fp32_inputs=False)
    got = ft.detach().reshape(-1)[band.idx.long()].cpu().numpy()
    ef = O.rel_l2(got, ref)
    g = _grad(band.n, 79).astype(np.complex128)
    box_g = torch.zeros(band.size, dtype=torch.complex128, device=device)
    box_g[band.idx.long()] = torch.tensor(g, device=device)
    (gv,) = torch.autograd.grad(ft, Vd, box_g.reshape(ft.shape))
    assert gv.dtype == torch.float64
    refg = O.acc_backward(Vs, w.faces, np.ones(len(w.faces)), xi64, g, threads=threads(),
                          fp32_inputs=False)
    eb = O.rel_l2(gv.cpu().numpy(), refg)
    # the native module's float64 entry points directly
    Vt = torch.tensor(Vs, device=device)
    xt = torch.tensor(xi64, device=device)
    ft2 = forward_fourier3d_cpp(Vt, Fc, C, xt)
    gv2 = backward_fourier3d_cpp(Vt, Fc, C, torch.tensor(g, device=device), ft2, xt)
    ef2, eb2 = O.rel_l2(ft2.cpu().numpy(), ref), O.rel_l2(gv2.cpu().numpy(), refg)
    print(f"float64 inputs, config 1: module forward {ef:.2e} adjoint {eb:.2e}; native forward "
          f"{ef2:.2e} adjoint {eb2:.2e}")
    assert max(ef, ef2) <= FWD_TOL and max(eb, eb2) <= BWD_TOL
    assert ft2.dtype == torch.complex128 and gv2.dtype == torch.float64


def test_full_size_cfg2_properties(device):
    """Config 2 at full size (all 1.08M band frequencies, 20,480 faces), checked
    through properties that need no oracle: Hermitian symmetry of the
    unmirrored forward, the translation phase exp(-i xi.t), invariance under
    uniform Faces_coeff scaling, linearity of the adjoint in the upstream
    gradient, and bitwise determinism."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config2()
    bs, axes, otf, band = _band(w, device)
    n = band.idx.numel()
    V = torch.tensor(_shifted(w, bs), device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    X = band.xi_list
    f = forward_fourier3d_cpp(V, Fc, C, X)
    fn = f.cpu().numpy().astype(np.complex128)
    scale = np.abs(fn).max()
    herm = np.abs(fn[::-1] - np.conj(fn)).max() / scale
    # a translation on a 2^-12 grid: exact in FP32 for every vertex whose sum
    # stays in its binade (all but a handful)
    tt = np.array([50.0, -89.0, 65.0]) / 4096.0
    t = torch.tensor(tt, dtype=torch.float32, device=device)
    ft = forward_fourier3d_cpp((V + t).contiguous(), Fc, C, X).cpu().numpy().astype(np.complex128)
    phase = np.exp(-1j * (X.cpu().numpy().astype(np.float64) @ tt))
    trans = O.rel_l2(ft, phase * fn)
    fs = forward_fourier3d_cpp(V, Fc, (2.5 * C).contiguous(), X)
    g1 = torch.tensor(_grad(n, 101), device=device)
    g2 = torch.tensor(_grad(n, 102), device=device)
    a = backward_fourier3d_cpp(V, Fc, C, g1, f, X)
    b = backward_fourier3d_cpp(V, Fc, C, g2, f, X)
    ab = backward_fourier3d_cpp(V, Fc, C, (g1 + g2).contiguous(), f, X)
    lin = O.rel_l2((a + b).cpu().numpy(), ab.cpu().numpy())
    f2 = forward_fourier3d_cpp(V, Fc, C, X)
    a2 = backward_fourier3d_cpp(V, Fc, C, g1, f, X)
    print(f"cfg2 full band n={n}: hermitian {herm:.2e}, translation rel-L2 {trans:.2e}, "
          f"adjoint linearity rel-L2 {lin:.2e}")
    assert herm <= 1e-6
    assert trans <= 1e-5
    assert torch.equal(fs.view(torch.float32), f.view(torch.float32))
    assert lin <= 1e-6
    assert torch.equal(f2.view(torch.float32), f.view(torch.float32)) and torch.equal(a2, a)


def test_far_translated_mesh_vertex_phases(device):
    """Vertex phases beyond 2^18 revolutions take the kernels' FP64 reduction
    instead of the fixed-point one: config 1's mesh (vertices rounded to
    multiples of 2^-8) translated by t ~ 1.2e4 on the same grid, so every
    translated vertex is exact in FP32, has |a/2pi| |xi| up to ~3e5
    revolutions on its band (frequencies below ~140 keep the fixed-point
    reduction: both forms run in one launch).  Checked against the FP64
    oracle on the translated inputs, against the untranslated forward times
    exp(-i xi.t), and the adjoint against the untranslated adjoint with the
    upstream gradient rotated by exp(-i xi.t) (the same function of the
    vertices)."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    V0 = np.round(_shifted(w, bs).astype(np.float64) * 256.0) / 256.0
    tt = np.array([8192.0, -8192.0, 4096.0]) + np.array([37.0, -91.0, 13.0]) / 256.0
    V1 = V0 + tt
    assert np.array_equal(V1.astype(np.float32).astype(np.float64), V1)
    X = band.xi_list
    xn = X.cpu().numpy().astype(np.float64)
    rev = np.abs(V1).max(axis=0) @ np.abs(xn).T / (2 * np.pi)
    assert rev.max() > 2.0 ** 18 > rev.min()
    Fc = torch.tensor(w.faces, device=device)
    C = torch.tensor(w.coeff, device=device)
    Vt0 = torch.tensor(V0, dtype=torch.float32, device=device)
    Vt1 = torch.tensor(V1, dtype=torch.float32, device=device)
    f0 = forward_fourier3d_cpp(Vt0, Fc, C, X)
    f1 = forward_fourier3d_cpp(Vt1, Fc, C, X)
    got1 = f1.cpu().numpy().astype(np.complex128)
    phase = np.exp(-1j * (xn @ tt))
    trans = O.rel_l2(got1, phase * f0.cpu().numpy().astype(np.complex128))
    ref1 = O.acc_forward(V1.astype(np.float32), w.faces, w.coeff, xn.astype(np.float32),
                         threads=threads())
    err1 = O.rel_l2(got1, ref1)
    g = _grad(band.idx.numel(), w.seed + 4)
    g0 = torch.tensor(g, device=device)
    g1 = torch.tensor((g * phase).astype(np.complex64), device=device)
    a0 = backward_fourier3d_cpp(Vt0, Fc, C, g0, f0, X).cpu().numpy().astype(np.float64)
    a1 = backward_fourier3d_cpp(Vt1, Fc, C, g1, f1, X).cpu().numpy().astype(np.float64)
    adj = O.rel_l2(a1, a0)
    print(f"far-translated mesh (max vertex phase {rev.max():.3g} revolutions): forward vs "
          f"oracle {err1:.2e}, vs phase x untranslated {trans:.2e}; adjoint vs untranslated "
          f"{adj:.2e}")
    assert err1 <= FWD_TOL and trans <= FWD_TOL
    assert adj <= BWD_TOL


def test_autofluorescence_forward_config1(device):
    """Autofluorescence variant (render.py:228-318) on config 1's full band
    (raw vertices, unit weights, regint 0.5) against the oracle, through the
    C-ABI with and without Hermitian halving."""
    from deltamic_amd.fourier3d_cpp import forward_fourier3d_cpp
    from deltamic_amd.hip_class import band_forward
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    V = torch.tensor(w.verts, device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.ones(len(w.faces), device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, band.xi_list, regint=0.5)
    xi = band.xi_list.cpu().numpy()
    ref = O.acc_forward_af(w.verts, w.faces, np.ones(len(w.faces)), 0.5, xi, threads=threads())
    err = O.rel_l2(ft.cpu().numpy(), ref)
    assert band.symmetric
    halved = band_forward(V, Fc, C, band, regint=0.5)
    herr = O.rel_l2(halved.cpu().numpy(), ft.cpu().numpy())
    print(f"autofluorescence cfg1 forward rel-L2 vs FP64 oracle: {err:.3e} (n={len(xi)}); "
          f"halved vs full {herr:.1e}")
    assert err <= FWD_TOL
    assert herr <= 2e-6


def test_autofluorescence_adjoint_vs_oracle_differences(device):
    """Gradient of sum Re(conj g ftmesh) for the autofluorescence variant
    (including the dependence of the normals and of the enclosed volume on the
    vertices) against 4th-order central differences of the FP64 oracle, on a
    jittered 42-vertex icosphere and 1024 frequencies."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    u, f = syn.icosphere(1)
    g = syn.rng(6101)
    v = (0.5 * u + g.normal(0, 0.01, u.shape) + 0.3).astype(np.float32)
    xi = g.normal(0, 6.0, (1024, 3)).astype(np.float32)
    gr = _grad(len(xi), 6102)
    regint = 0.6
    V = torch.tensor(v, device=device)
    Fc = torch.tensor(f, device=device)
    C = torch.ones(len(f), device=device)
    X = torch.tensor(xi, device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, X, regint=regint)
    gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(gr, device=device), ft, X,
                                regint=regint).cpu().numpy().astype(np.float64)
    v64 = v.astype(np.float64)
    ones = np.ones(len(f))

    def L(vv):
        a = O.acc_forward_af(vv, f, ones, regint, xi, threads=threads(), fp32_inputs=False)
        return float(np.sum(gr.real * a.real + gr.imag * a.imag))

    h = 1e-4
    fd = np.zeros_like(v64)
    for i in range(len(v64)):
        for d in range(3):
            vals = []
            for k in (-2, -1, 1, 2):
                vv = v64.copy()
                vv[i, d] += k * h
                vals.append(L(vv))
            fd[i, d] = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
    err = np.linalg.norm(gv - fd) / np.linalg.norm(fd)
    print(f"autofluorescence adjoint rel-L2 vs oracle differences: {err:.3e}")
    assert err <= BWD_TOL


def test_autofluorescence_module_function(device):
    """fourier_from_mesh_narrowband_autofluorescence: band values equal the
    C-ABI call, zeros outside the band, and autograd reaches the vertices with
    the adjoint of the C-ABI."""
    from deltamic_amd import fourier_from_mesh_narrowband_autofluorescence
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    bs, axes, otf, band = _band(w, device)
    V = torch.tensor(w.verts, device=device, requires_grad=True)
    Fc = torch.tensor(w.faces, device=device)
    box = fourier_from_mesh_narrowband_autofluorescence(otf, w.thresh, 0.5, *axes, V, Fc)
    assert box.shape == tuple(w.box_shape) and box.dtype == torch.complex64
    flat = box.detach().reshape(-1)
    vals = flat[band.idx.long()]
    C = torch.ones(len(w.faces), device=device)
    direct = forward_fourier3d_cpp(V.detach(), Fc, C, band.xi_list, regint=0.5)
    assert O.rel_l2(vals.cpu().numpy(), direct.cpu().numpy()) <= 2e-6
    assert int((flat != 0).sum()) <= band.n
    up = torch.tensor(_grad(int(np.prod(w.box_shape)), 6103).reshape(w.box_shape), device=device)
    (box.real * up.real + box.imag * up.imag).sum().backward()
    g_list = up.reshape(-1)[band.idx.long()].contiguous()
    ref = backward_fourier3d_cpp(V.detach(), Fc, C, g_list, direct, band.xi_list, regint=0.5)
    err = O.rel_l2(V.grad.cpu().numpy(), ref.cpu().numpy())
    print(f"autofluorescence module gradient vs direct adjoint: {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("nf,n", [(1, 1), (63, 7), (64, 8), (65, 9), (65, 300), (700, 5000),
                                  (130, 2049)])
def test_work_decomposition_sizes(device, nf, n):
    """Forward and adjoint against the FP64 oracle at sizes on and off the
    kernels' block boundaries: 64-face blocks (one lane per face), 8-frequency
    forward workgroups, 64-frequency adjoint tiles and the chunk plans (equal
    chunks for small bands, geometric levels per XCD eighth for large ones)."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    r = syn.rng(7000 + nf * 13 + n)
    nv = max(3, nf // 2 + 2)
    V = (0.3 * r.standard_normal((nv, 3)) + 1.0).astype(np.float32)
    Fc = np.stack([r.permutation(nv)[:3] for _ in range(nf)]).astype(np.int64)
    C = (0.5 + r.random(nf)).astype(np.float32)
    xi = (6.0 * r.standard_normal((n, 3))).astype(np.float32)
    g = (r.standard_normal(n) + 1j * r.standard_normal(n)).astype(np.complex64)
    Vd, Fd, Cd = (torch.tensor(a, device=device) for a in (V, Fc, C))
    xid = torch.tensor(xi, device=device)
    ft = forward_fourier3d_cpp(Vd, Fd, Cd, xid)
    ref = O.acc_forward(V, Fc, C, xi)
    ferr = O.rel_l2(ft.cpu().numpy(), ref)
    gv = backward_fourier3d_cpp(Vd, Fd, Cd, torch.tensor(g, device=device), ft, xid)
    gref = O.acc_backward(V, Fc, C, xi, g)
    berr = O.rel_l2(gv.cpu().numpy(), gref)
    print(f"nf={nf} n={n}: forward rel-L2 {ferr:.2e}, adjoint rel-L2 {berr:.2e}")
    assert ferr <= FWD_TOL and berr <= BWD_TOL
