"""Gaussian OTF on the meshFT frequency grid (drop-in for deltamic/psf.py:15-20)."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _gaussian_psf_torch(sigma, xi0, xi1, xi2):
    # psf.py:16-19, kept for a differentiable sigma: autograd needs torch ops
    g0, g1, g2 = torch.meshgrid(xi0, xi1, xi2, indexing='ij')
    return torch.exp(-(g0 * (sigma[0, 0] * g0 + sigma[0, 1] * g1 + sigma[0, 2] * g2)
                       + g1 * (sigma[1, 0] * g0 + sigma[1, 1] * g1 + sigma[1, 2] * g2)
                       + g2 * (sigma[2, 0] * g0 + sigma[2, 1] * g1 + sigma[2, 2] * g2)) / 2)


def generate_gaussian_psf(sigma, xi0, xi1, xi2):
    """OTF[i0, i1, i2] = exp(-xi^T Sigma xi / 2) with xi = (xi0[i0], xi1[i1], xi2[i2]).

    On GPU float32 axes with a constant ``sigma`` this is one HIP kernel that
    writes the box directly (no three n0*n1*n2 meshgrid temporaries).  A
    ``sigma`` that requires grad keeps the reference's torch expression so
    autograd reaches it; that is the learnable-PSF case, outside the meshFT
    kernels.
    """
    sig_t = sigma if isinstance(sigma, torch.Tensor) else torch.as_tensor(np.asarray(sigma))
    needs_grad = sig_t.requires_grad and torch.is_grad_enabled()
    on_gpu = all(isinstance(a, torch.Tensor) and a.is_cuda for a in (xi0, xi1, xi2))
    if needs_grad or not on_gpu or xi0.dtype != torch.float32:
        return _gaussian_psf_torch(sig_t.to(xi0.device), xi0, xi1, xi2)
    s = np.ascontiguousarray(sig_t.detach().cpu().numpy().reshape(9), dtype=np.float32)
    axes = [a.detach().to(torch.float32).contiguous() for a in (xi0, xi1, xi2)]
    n0, n1, n2 = (a.numel() for a in axes)
    out = torch.empty((n0, n1, n2), dtype=torch.float32, device=axes[0].device)
    L = _lib.lib()
    _lib.check(L.dmic_gaussian_otf(s.ctypes.data, axes[0].data_ptr(), axes[1].data_ptr(),
                                   axes[2].data_ptr(), n0, n1, n2, out.data_ptr(),
                                   _lib.stream_of(out)), "dmic_gaussian_otf")
    return out
