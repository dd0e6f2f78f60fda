"""GPU numerics of the learnable PSF kernels (csrc/psf.hip) and of the
learnable-PSF fit step (PSF -> OTF -> band -> meshFT -> render -> loss).

References: the formulas of deltamic/psf.py:15-20,551-665 evaluated with
torch on the CPU in float64, differentiated by torch autograd.  Tolerances:
PSF values 2e-7 (1 + exponent) relative (FP32 exponent and expf), parameter
gradients 1e-5 rel-L2 (FP32 products, FP64 sums), the gradient of the whole
fit step w.r.t. the PSF matrix 1e-4 (FP32 FFTs).
"""
import numpy as np
import pytest
import torch

from deltamic_amd import synthetic as syn

pytestmark = pytest.mark.gpu

SHAPE = (40, 48, 56)
BOX = torch.tensor([1.3, 1.1, 1.7], dtype=torch.float64)


def _ref_gauss(S, g0, g1, g2):
    # psf.py:573-576
    return torch.exp(-(g0 * (S[0, 0] * g0 + S[0, 1] * g1 + S[0, 2] * g2)
                       + g1 * (S[1, 0] * g0 + S[1, 1] * g1 + S[1, 2] * g2)
                       + g2 * (S[2, 0] * g0 + S[2, 1] * g1 + S[2, 2] * g2)) / 2)


def _ref_aniso(S, sz, eps, g0, g1, g2):
    # psf.py:656-658
    D = eps + (torch.abs(g2) * sz) ** 2
    return torch.exp(-(g0 * (S[0, 0] * g0 + S[0, 1] * g1) + g1 * (S[1, 0] * g0 + S[1, 1] * g1))
                     / (2 * D)) / D


def _grids64(m):
    return torch.meshgrid(m.x0.cpu().double(), m.x1.cpu().double(), m.x2.cpu().double(),
                          indexing='ij')


def _check_values(got, want, q):
    got = got.detach().cpu().double()
    live = want > 1e-30
    rel = (got[live] - want[live]).abs() / want[live]
    assert (rel <= 2e-7 * (4.0 + q[live].abs())).all(), float(rel.max())


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_gaussian_psf_module_forward_and_sigma_gradient(device):
    from deltamic_amd import Gaussian_psf
    m = Gaussian_psf(SHAPE, BOX, sigma=30.0, device=device)
    S0 = torch.tensor([[30.0, 4.0, -2.0], [3.0, 45.0, 5.0], [-1.0, 6.0, 20.0]])
    with torch.no_grad():
        m.sigma.copy_(S0)
    W = torch.tensor(syn.rng(11).normal(size=SHAPE), dtype=torch.float32, device=device)
    psf = m()
    (psf * W).sum().backward()
    S = S0.double().requires_grad_(True)
    g0, g1, g2 = _grids64(m)
    want = _ref_gauss(S, g0, g1, g2)
    q = -2.0 * torch.log(want.detach().clamp_min(1e-300))
    _check_values(psf, want.detach(), q)
    (want * W.cpu().double()).sum().backward()
    err = _rel(m.sigma.grad.cpu().double(), S.grad)
    print(f"Gaussian_psf: sigma gradient rel-L2 {err:.2e}")
    assert err <= 1e-5


def test_gaussian_psf_logsigma_chain(device):
    from deltamic_amd import Gaussian_psf_logsigma
    m = Gaussian_psf_logsigma(SHAPE, BOX, sigma=25.0, device=device)
    W = torch.tensor(syn.rng(12).normal(size=SHAPE), dtype=torch.float32, device=device)
    (m() * W).sum().backward()
    L = (torch.eye(3) * np.log(25.0)).double().requires_grad_(True)
    g0, g1, g2 = _grids64(m)
    (_ref_gauss(torch.exp(L), g0, g1, g2) * W.cpu().double()).sum().backward()
    err = _rel(m.logsigma.grad.cpu().double(), L.grad)
    print(f"Gaussian_psf_logsigma: logsigma gradient rel-L2 {err:.2e}")
    assert err <= 1e-5


def test_anisotropic_psf_forward_and_gradients(device):
    from deltamic_amd import Anisotropic_gaussian_psf
    m = Anisotropic_gaussian_psf(SHAPE, BOX, sigma=8.0, sigma_z=1.5, eps=0.05, device=device)
    with torch.no_grad():
        m.sigma.copy_(torch.tensor([[8.0, 1.5], [0.5, 11.0]]))
    W = torch.tensor(syn.rng(13).normal(size=SHAPE), dtype=torch.float32, device=device)
    psf = m()
    (psf * W).sum().backward()
    S = torch.tensor([[8.0, 1.5], [0.5, 11.0]], dtype=torch.float64, requires_grad=True)
    sz = torch.tensor(1.5, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    g0, g1, g2 = _grids64(m)
    want = _ref_aniso(S, sz, eps, g0, g1, g2)
    got = psf.detach().cpu().double()
    assert _rel(got, want.detach()) <= 1e-6
    (want * W.cpu().double()).sum().backward()
    errs = [_rel(m.sigma.grad.cpu().double(), S.grad),
            abs(float(m.sigma_z.grad) - float(sz.grad)) / abs(float(sz.grad)),
            abs(float(m.eps.grad) - float(eps.grad)) / abs(float(eps.grad))]
    print(f"Anisotropic_gaussian_psf: gradient errors sigma {errs[0]:.2e}, sigma_z "
          f"{errs[1]:.2e}, eps {errs[2]:.2e}")
    assert max(errs) <= 1e-5


def test_generate_gaussian_psf_differentiable_sigma(device):
    """generate_gaussian_psf with a sigma that requires grad (the OTF on the
    frequency grid, psf.py:15-20) goes through the HIP autograd Function."""
    from deltamic_amd import generate_gaussian_psf, spatial_frequency_axes
    w = syn.config1()
    axes = spatial_frequency_axes(torch.tensor(w.box_size, dtype=torch.float32, device=device),
                                  w.box_shape, torch.float32, device)
    S = torch.tensor(w.sigma, dtype=torch.float32, device=device).requires_grad_(True)
    W = torch.tensor(syn.rng(14).normal(size=tuple(w.box_shape)), dtype=torch.float32,
                     device=device)
    otf = generate_gaussian_psf(S, *axes)
    (otf * W).sum().backward()
    S64 = torch.tensor(w.sigma, dtype=torch.float64, requires_grad=True)
    g = torch.meshgrid(*[a.cpu().double() for a in axes], indexing='ij')
    want = _ref_gauss(S64, *g)
    assert _rel(otf.detach().cpu().double(), want.detach()) <= 1e-6
    (want * W.cpu().double()).sum().backward()
    err = _rel(S.grad.cpu().double(), S64.grad)
    print(f"generate_gaussian_psf: sigma gradient rel-L2 {err:.2e}")
    assert err <= 1e-5


def test_learnable_psf_fit_step(device):
    """One step of the reference's fit loop with a learnable PSF
    (Examples/Gradient_Descent.ipynb cell 2): PSF -> OTF = |fftshift(fftn(PSF))|
    -> meshFT.OTF (band recompacted) -> render -> MSE.  The PSF-matrix gradient
    is checked against the same chain in float64 on the CPU with the band
    values of the GPU forward held fixed; the band follows the OTF."""
    from deltamic_amd import Fourier3dMesh, Gaussian_psf, normalize_tensor, render_image_from_ftmesh
    w = syn.config1()
    box = torch.tensor(w.box_size)
    shape = tuple(w.box_shape)
    meshFT = Fourier3dMesh(box, shape, device=device, dtype=torch.float32, narrowband_thresh=1e-2)
    model = Gaussian_psf(shape, box[:, 1], sigma=1e3, device=device)
    V = torch.tensor(w.verts, device=device, requires_grad=True)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.ones(len(w.faces), device=device)
    target = torch.tensor(syn.rng(15).random(size=shape), dtype=torch.float32, device=device)
    bands = []
    for it in range(2):
        if it == 1:
            with torch.no_grad():
                model.sigma.mul_(0.8)  # a new PSF must give a new band
        model.zero_grad()
        V.grad = None
        PSF = model()
        OTF = torch.abs(torch.fft.fftshift(torch.fft.fftn(PSF)))
        meshFT.OTF = OTF
        ftmesh = meshFT(V, Fc, C)
        img = normalize_tensor(render_image_from_ftmesh(ftmesh, OTF, shape))
        loss = torch.mean((img - target) ** 2)
        loss.backward()
        band = int((OTF.detach() > 1e-2).sum())
        bands.append(band)
        assert torch.isfinite(V.grad).all() and float(V.grad.abs().max()) > 0
        # float64 CPU restatement of the PSF -> OTF -> render -> loss chain
        S = model.sigma.detach().cpu().double().requires_grad_(True)
        g = torch.meshgrid(model.x0.cpu().double(), model.x1.cpu().double(),
                           model.x2.cpu().double(), indexing='ij')
        otf64 = torch.abs(torch.fft.fftshift(torch.fft.fftn(_ref_gauss(S, *g))))
        ft64 = ftmesh.detach().cpu().to(torch.complex128)
        img64 = normalize_tensor(torch.abs(torch.fft.ifftn(ft64 * otf64, shape)))
        torch.mean((img64 - target.cpu().double()) ** 2).backward()
        err = _rel(model.sigma.grad.cpu().double(), S.grad)
        print(f"learnable-PSF step {it}: band {band}, sigma gradient rel-L2 {err:.2e}")
        assert err <= 1e-4
    assert bands[0] != bands[1]
