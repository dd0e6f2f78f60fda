"""``torch.autograd.Function`` of the GPU meshFT path.

Mirrors ``MeshFTCUDA`` of deltamic/cuda_class.py:10-40 (same arguments, same
saved state, same gradient return ``(gradVerts, None x 7)``), with the band
work done by HIP kernels:

* band compaction ``Filter > thresh`` (no abs, cuda_class.py:21) in
  row-major voxel order without materialising a meshgrid (cuda_class.py:20-23);
* the band is cached on the ``Filter`` tensor and rebuilt only when that
  tensor, its version counter, the threshold or the frequency axes change
  (the reference rebuilds it on every forward);
* scatter into the full complex box (cuda_class.py:27-28) and the gather of
  ``grad_output`` (cuda_class.py:36).

Unlike cuda_class.py:8 importing this module does not reseed torch's RNG.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch
from torch.autograd import Function

from . import _lib
from .fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp


@dataclass
class Band:
    idx: torch.Tensor       # int32 [n] linear voxel indices, row-major
    xi_list: torch.Tensor   # float32 [n, 3]
    shape: tuple
    key: tuple


def _thresh_value(narrowband_thresh) -> float:
    if isinstance(narrowband_thresh, torch.Tensor):
        return float(narrowband_thresh.detach().float().cpu())
    return float(narrowband_thresh)


def compute_band(Filter: torch.Tensor, thresh: float, xi0: torch.Tensor, xi1: torch.Tensor,
                 xi2: torch.Tensor) -> Band:
    """Compacts the band ``Filter > thresh`` (HIP kernels; one host sync for the count)."""
    if not Filter.is_cuda:
        raise RuntimeError("Filter must be a CUDA tensor")
    if Filter.dim() != 3:
        raise ValueError(f"Filter must be 3-D, got {tuple(Filter.shape)}")
    n0, n1, n2 = Filter.shape
    f = Filter.detach()
    if f.dtype != torch.float32:
        f = f.float()
    f = f.contiguous()
    axes = []
    for a, nd in ((xi0, n0), (xi1, n1), (xi2, n2)):
        if a.numel() != nd:
            raise ValueError("frequency axis length does not match Filter shape")
        axes.append(a.detach().to(device=f.device, dtype=torch.float32).contiguous())
    dev = f.device
    L = _lib.lib()
    st = _lib.stream_of(f)
    ws = _lib.workspace(L.dmic_band_workspace(n0, n1, n2), dev)
    count = torch.zeros((1,), dtype=torch.int64, device=dev)
    _lib.check(L.dmic_band_count(f.data_ptr(), n0, n1, n2, float(thresh), count.data_ptr(),
                                 ws.data_ptr(), ws.numel(), st), "dmic_band_count")
    n = int(count.item())
    idx = torch.empty((n,), dtype=torch.int32, device=dev)
    xi_list = torch.empty((n, 3), dtype=torch.float32, device=dev)
    if n:
        _lib.check(L.dmic_band_fill(f.data_ptr(), n0, n1, n2, float(thresh), axes[0].data_ptr(),
                                    axes[1].data_ptr(), axes[2].data_ptr(), idx.data_ptr(),
                                    xi_list.data_ptr(), ws.data_ptr(), ws.numel(), st),
                   "dmic_band_fill")
    return Band(idx=idx, xi_list=xi_list, shape=(n0, n1, n2), key=())


def band_for(Filter: torch.Tensor, narrowband_thresh, xi0, xi1, xi2) -> Band:
    """Band of ``Filter``, cached on the tensor object itself."""
    thresh = _thresh_value(narrowband_thresh)
    key = (Filter._version, thresh, id(xi0), xi0._version, id(xi1), xi1._version, id(xi2),
           xi2._version)
    cached = getattr(Filter, "_dmic_band", None)
    if cached is not None and cached[0] == key and cached[2] is xi0 and cached[3] is xi1 \
            and cached[4] is xi2:
        return cached[1]
    band = compute_band(Filter, thresh, xi0, xi1, xi2)
    band.key = key
    Filter._dmic_band = (key, band, xi0, xi1, xi2)
    return band


def scatter_band(vals: torch.Tensor, band: Band) -> torch.Tensor:
    box = torch.empty(band.shape, dtype=torch.complex64, device=vals.device)
    L = _lib.lib()
    _lib.check(L.dmic_band_scatter(vals.data_ptr(), band.idx.data_ptr(), band.idx.numel(),
                                   box.data_ptr(), box.numel(), _lib.stream_of(box)),
               "dmic_band_scatter")
    return box


def gather_band(box: torch.Tensor, band: Band) -> torch.Tensor:
    box = box.contiguous()
    if box.dtype != torch.complex64:
        box = box.to(torch.complex64)
    vals = torch.empty((band.idx.numel(),), dtype=torch.complex64, device=box.device)
    L = _lib.lib()
    _lib.check(L.dmic_band_gather(box.data_ptr(), band.idx.data_ptr(), band.idx.numel(),
                                  vals.data_ptr(), _lib.stream_of(box)), "dmic_band_gather")
    return vals


class MeshFTHIP(Function):
    """HIP-based Function for the Fourier transform of a triangle mesh."""

    @staticmethod
    def forward(ctx, Verts, Faces, Faces_coeff, xi0, xi1, xi2, Filter, narrowband_thresh):
        if Verts.dtype != torch.float32:
            raise TypeError("Mesh FT (HIP) is implemented for float32 only, got " + str(Verts.dtype))
        band = band_for(Filter, narrowband_thresh, xi0, xi1, xi2)
        V = Verts.detach().contiguous()
        Fc = Faces.contiguous()
        C = Faces_coeff.detach().to(torch.float32).contiguous()
        ftmesh_list = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
        ctx.save_for_backward(V, Fc, C, ftmesh_list)
        ctx.band = band
        ftmesh = scatter_band(ftmesh_list, band)
        if Filter.dtype == torch.float64:  # output dtype follows Filter (cuda_class.py:27)
            ftmesh = ftmesh.to(torch.complex128)
        return ftmesh

    @staticmethod
    def backward(ctx, grad_output):
        V, Fc, C, ftmesh_list = ctx.saved_tensors
        band = ctx.band
        grad_list = gather_band(grad_output, band)
        gradVerts = backward_fourier3d_cpp(V, Fc, C, grad_list, ftmesh_list, band.xi_list)
        return gradVerts, None, None, None, None, None, None, None


# name of the reference's Function, for code that imports it directly
MeshFTCUDA = MeshFTHIP
