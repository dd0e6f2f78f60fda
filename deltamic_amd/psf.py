"""Gaussian OTF on the meshFT frequency grid (drop-in for deltamic/psf.py:15-20)
and the learnable Gaussian PSF modules of deltamic/psf.py:551-665
(``Gaussian_psf``, ``Gaussian_psf_logsigma``, ``Anisotropic_gaussian_psf``).

Every evaluation is a HIP kernel (csrc/band.hip, csrc/psf.hip); a parameter
that requires grad goes through an autograd Function whose backward is the
HIP reduction of the parameter gradient.  There is no CPU path: CPU tensors
raise, as the meshFT module does.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import nn

from . import _lib


def _gpu_axes(xi0, xi1, xi2):
    axes = (xi0, xi1, xi2)
    if not all(isinstance(a, torch.Tensor) and a.is_cuda for a in axes):
        raise NotImplementedError("deltamic_amd evaluates PSFs on the GPU only; "
                                  "pass CUDA (HIP) tensors")
    if any(a.dtype != torch.float32 for a in axes):
        raise TypeError("PSF axes must be float32 (the HIP kernels are FP32)")
    return [a.detach().contiguous() for a in axes]


class _GaussianPSF(torch.autograd.Function):
    """psf = exp(-x^T S x / 2) on the grid of three axes; gradient w.r.t. S."""

    @staticmethod
    def forward(ctx, S, x0, x1, x2):
        if S.numel() != 9:  # the kernel reads a 3 x 3 matrix
            raise ValueError(f"sigma must be 3 x 3, got {tuple(S.shape)}")
        axes = _gpu_axes(x0, x1, x2)
        Sd = S.detach().to(device=axes[0].device, dtype=torch.float32).contiguous()
        n0, n1, n2 = (a.numel() for a in axes)
        out = torch.empty((n0, n1, n2), dtype=torch.float32, device=axes[0].device)
        L = _lib.lib()
        _lib.check(L.dmic_psf_gaussian(Sd.data_ptr(), axes[0].data_ptr(), axes[1].data_ptr(),
                                       axes[2].data_ptr(), n0, n1, n2, out.data_ptr(),
                                       _lib.stream_of(out)), "dmic_psf_gaussian")
        ctx.save_for_backward(out, *axes)
        ctx.s_dtype, ctx.s_device = S.dtype, S.device
        return out

    @staticmethod
    def backward(ctx, grad):
        out, x0, x1, x2 = ctx.saved_tensors
        if not ctx.needs_input_grad[0]:
            return None, None, None, None
        n0, n1, n2 = out.shape
        g = grad.to(torch.float32).contiguous()
        gS = torch.empty(9, dtype=torch.float32, device=out.device)
        nbytes = int(_lib.lib().dmic_psf_grad_workspace(n0, n1, n2))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=out.device)
        _lib.check(_lib.lib().dmic_psf_gaussian_grad(
            g.data_ptr(), out.data_ptr(), x0.data_ptr(), x1.data_ptr(), x2.data_ptr(), n0, n1, n2,
            gS.data_ptr(), ws.data_ptr(), nbytes, _lib.stream_of(out)), "dmic_psf_gaussian_grad")
        return gS.view(3, 3).to(dtype=ctx.s_dtype, device=ctx.s_device), None, None, None


class _AnisotropicPSF(torch.autograd.Function):
    """psf = exp(-Q/(2D))/D, prm = (S2_00, S2_01, S2_10, S2_11, sz, eps)."""

    @staticmethod
    def forward(ctx, prm, x0, x1, x2):
        if prm.numel() != 6:  # the kernel reads six parameters
            raise ValueError(f"anisotropic PSF takes 6 parameters, got {prm.numel()}")
        axes = _gpu_axes(x0, x1, x2)
        p = prm.detach().to(device=axes[0].device, dtype=torch.float32).contiguous()
        n0, n1, n2 = (a.numel() for a in axes)
        out = torch.empty((n0, n1, n2), dtype=torch.float32, device=axes[0].device)
        L = _lib.lib()
        _lib.check(L.dmic_psf_aniso(p.data_ptr(), axes[0].data_ptr(), axes[1].data_ptr(),
                                    axes[2].data_ptr(), n0, n1, n2, out.data_ptr(),
                                    _lib.stream_of(out)), "dmic_psf_aniso")
        ctx.save_for_backward(p, out, *axes)
        ctx.p_dtype, ctx.p_device = prm.dtype, prm.device
        return out

    @staticmethod
    def backward(ctx, grad):
        p, out, x0, x1, x2 = ctx.saved_tensors
        if not ctx.needs_input_grad[0]:
            return None, None, None, None
        n0, n1, n2 = out.shape
        g = grad.to(torch.float32).contiguous()
        gp = torch.empty(6, dtype=torch.float32, device=out.device)
        nbytes = int(_lib.lib().dmic_psf_grad_workspace(n0, n1, n2))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=out.device)
        _lib.check(_lib.lib().dmic_psf_aniso_grad(
            p.data_ptr(), g.data_ptr(), out.data_ptr(), x0.data_ptr(), x1.data_ptr(),
            x2.data_ptr(), n0, n1, n2, gp.data_ptr(), ws.data_ptr(), nbytes,
            _lib.stream_of(out)), "dmic_psf_aniso_grad")
        return gp.to(dtype=ctx.p_dtype, device=ctx.p_device), None, None, None


def generate_gaussian_psf(sigma, xi0, xi1, xi2):
    """OTF[i0, i1, i2] = exp(-xi^T Sigma xi / 2) with xi = (xi0[i0], xi1[i1], xi2[i2])
    (psf.py:15-20).

    One HIP kernel writes the box directly (no three n0*n1*n2 meshgrid
    temporaries).  A ``sigma`` that requires grad goes through an autograd
    Function whose backward reduces dL/dSigma on the GPU (the learnable-PSF
    case of the reference's fit loop)."""
    sig_t = sigma if isinstance(sigma, torch.Tensor) else torch.as_tensor(np.asarray(sigma))
    if sig_t.numel() != 9:
        raise ValueError(f"sigma must be 3 x 3, got {tuple(sig_t.shape)}")
    if all(isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == torch.float64
           for a in (xi0, xi1, xi2)):
        # float64 axes (a float64 Fourier3dMesh): the reference's formula in
        # FP64 torch ops on the GPU, separable in the axes; the HIP kernel is FP32
        S = sig_t.to(device=xi0.device, dtype=torch.float64).reshape(3, 3)
        a0, a1, a2 = xi0[:, None, None], xi1[None, :, None], xi2[None, None, :]
        q = (a0 * (S[0, 0] * a0 + S[0, 1] * a1 + S[0, 2] * a2)
             + a1 * (S[1, 0] * a0 + S[1, 1] * a1 + S[1, 2] * a2)
             + a2 * (S[2, 0] * a0 + S[2, 1] * a1 + S[2, 2] * a2))
        return torch.exp(-q / 2)
    axes = _gpu_axes(xi0, xi1, xi2)
    if sig_t.requires_grad and torch.is_grad_enabled():
        return _GaussianPSF.apply(sig_t, *axes)
    s = np.ascontiguousarray(sig_t.detach().cpu().numpy().reshape(9), dtype=np.float32)
    n0, n1, n2 = (a.numel() for a in axes)
    out = torch.empty((n0, n1, n2), dtype=torch.float32, device=axes[0].device)
    L = _lib.lib()
    _lib.check(L.dmic_gaussian_otf(s.ctypes.data, axes[0].data_ptr(), axes[1].data_ptr(),
                                   axes[2].data_ptr(), n0, n1, n2, out.data_ptr(),
                                   _lib.stream_of(out)), "dmic_gaussian_otf")
    return out


def _axes(box_shape, box_size, dtype, device):
    # psf.py:570-572: torch.linspace(-L/2, L/2, n) per axis
    return [torch.linspace(-float(box_size[d]) / 2, float(box_size[d]) / 2, int(box_shape[d]),
                           dtype=dtype, device=device) for d in range(3)]


class _PSFModule(nn.Module):
    def _init_grid(self, box_shape, box_size, device, dtype):
        self.x0, self.x1, self.x2 = _axes(box_shape, box_size, dtype, device)
        # the reference keeps full meshgrids; expanded views cost no memory here
        self.Gridxi0, self.Gridxi1, self.Gridxi2 = torch.meshgrid(self.x0, self.x1, self.x2,
                                                                  indexing='ij')

    def make_param(self, x):
        return torch.nn.parameter.Parameter(torch.tensor(x, dtype=self.dtype,
                                                         requires_grad=self.with_grad,
                                                         device=self.device))


class Gaussian_psf(_PSFModule):
    """Learnable Gaussian PSF with a full 3x3 matrix ``sigma`` (psf.py:551-586):
    forward() = exp(-x^T sigma x / 2) on the box grid."""

    def __init__(self, box_shape, box_size, sigma=1, device='cpu', with_grad=True,
                 dtype=torch.float):
        super().__init__()
        self.with_grad = with_grad
        self.dtype = dtype
        self.device = device
        self._init_grid(box_shape, box_size, device, dtype)
        self.sigma = self.make_param(np.identity(3) * sigma)

    def forward(self):
        return _GaussianPSF.apply(self.sigma, self.x0, self.x1, self.x2)


class Gaussian_psf_logsigma(_PSFModule):
    """Gaussian PSF parametrised by ``logsigma`` (psf.py:589-625): sigma =
    exp(logsigma) elementwise, so the off-diagonal entries start at exp(0) = 1
    as in the reference."""

    def __init__(self, box_shape, box_size, sigma=1, device='cpu', with_grad=True,
                 dtype=torch.float):
        super().__init__()
        self.with_grad = with_grad
        self.dtype = dtype
        self.device = device
        self._init_grid(box_shape, box_size, device, torch.float)  # psf.py:609-611: default dtype
        self.logsigma = self.make_param(np.identity(3) * np.log(sigma))

    def forward(self):
        self.sigma = torch.exp(self.logsigma)
        return _GaussianPSF.apply(self.sigma, self.x0, self.x1, self.x2)


class Anisotropic_gaussian_psf(_PSFModule):
    """Depth-dependent Gaussian PSF (psf.py:628-665): exp(-Q/(2D))/D with
    Q = (x0, x1) sigma (x0, x1)^T and D = eps + (|x2| sigma_z)^2."""

    def __init__(self, box_shape, box_size, sigma=1, sigma_z=1, eps=0.01, device='cpu',
                 with_grad=True, dtype=torch.float):
        super().__init__()
        self.with_grad = with_grad
        self.dtype = dtype
        self.device = device
        self._init_grid(box_shape, box_size, device, torch.float)
        self.sigma = self.make_param(np.identity(2) * sigma)
        self.sigma_z = self.make_param(sigma_z)
        self.eps = self.make_param(eps)

    def forward(self):
        prm = torch.cat([self.sigma.reshape(4), self.sigma_z.reshape(1), self.eps.reshape(1)])
        return _AnisotropicPSF.apply(prm, self.x0, self.x1, self.x2)
