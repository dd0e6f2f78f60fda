"""Module-level API of the meshFT path (drop-in for deltamic/render.py:5-180).

``Fourier3dMesh``, ``render_image_from_ftmesh``, ``normalize_tensor`` and
``compute_box_size`` keep the reference's names, argument order, defaults and
semantics.  On a GPU device the transform runs as HIP kernels (hip_class.py);
``device='cpu'`` is the reference's CPU device (``MeshFTPy``, render.py:36-109)
rebuilt in cpu_path.py with torch on the host.  A GPU module never falls back
to the CPU path: its boundary raises when its inputs are not on the GPU
(fourier3d_cpp.py, as fourier3d_cuda.cpp:23-25).
"""
from __future__ import annotations

import numpy as np
import torch
from torch import nn

from . import _lib
from .cpu_path import meshft_af_cpu, meshft_cpu
from .hip_class import MeshFTAutofluorescenceHIP, MeshFTHIP, clear_band_cache


class _NormalizeHIP(torch.autograd.Function):
    """normalize_tensor on a CUDA float32 tensor with the HIP kernels of
    csrc/render.hip: a min/max pass (skipped when the tensor comes from
    render_image_from_ftmesh, whose abs pass already reduced it) and one
    elementwise pass; backward is one reduction and one elementwise pass,
    with torch's gradient through min and max (shared evenly by ties)."""

    @staticmethod
    def forward(ctx, t, c):
        L = _lib.lib()
        st = _lib.stream_of(t)
        ws = _lib.workspace(L.dmic_render_workspace(t.numel()), t.device)
        cached = getattr(t, "_dmic_minmax", None)
        if cached is not None and cached[1] == t._version:
            mm = cached[0]
        else:
            mm = torch.empty(2, dtype=torch.float32, device=t.device)
            _lib.check(L.dmic_minmax(t.data_ptr(), t.numel(), mm.data_ptr(), ws.data_ptr(),
                                     ws.numel(), st), "dmic_minmax")
        out = torch.empty_like(t)
        _lib.check(L.dmic_normalize(t.data_ptr(), t.numel(), mm.data_ptr(), c, out.data_ptr(), st),
                   "dmic_normalize")
        ctx.save_for_backward(t, mm)
        ctx.c = c
        return out

    @staticmethod
    def backward(ctx, g):
        t, mm = ctx.saved_tensors
        g = g.contiguous().to(torch.float32)
        L = _lib.lib()
        ws = _lib.workspace(L.dmic_render_workspace(t.numel()), t.device)
        grad_t = torch.empty_like(t)
        _lib.check(L.dmic_normalize_backward(g.data_ptr(), t.data_ptr(), t.numel(), mm.data_ptr(),
                                             ctx.c, grad_t.data_ptr(), ws.data_ptr(), ws.numel(),
                                             _lib.stream_of(t)), "dmic_normalize_backward")
        return grad_t, None


class _AbsHIP(torch.autograd.Function):
    """|z| of a CUDA complex64 box (HIP kernel; its min and max come out of
    the same pass and are kept on the result for normalize_tensor);
    backward g z / |z| (torch's sgn, 0 at z = 0)."""

    @staticmethod
    def forward(ctx, z):
        L = _lib.lib()
        st = _lib.stream_of(z)
        img = torch.empty(z.shape, dtype=torch.float32, device=z.device)
        mm = torch.empty(2, dtype=torch.float32, device=z.device)
        ws = _lib.workspace(L.dmic_render_workspace(z.numel()), z.device)
        _lib.check(L.dmic_abs_minmax(z.data_ptr(), z.numel(), img.data_ptr(), mm.data_ptr(),
                                     ws.data_ptr(), ws.numel(), st), "dmic_abs_minmax")
        ctx.save_for_backward(z)
        img._dmic_mm = mm
        return img

    @staticmethod
    def backward(ctx, g):
        (z,) = ctx.saved_tensors
        g = g.contiguous().to(torch.float32)
        gz = torch.empty_like(z)
        L = _lib.lib()
        _lib.check(L.dmic_abs_backward(g.data_ptr(), z.data_ptr(), z.numel(), gz.data_ptr(),
                                       _lib.stream_of(z)), "dmic_abs_backward")
        return gz


def _hip_ok(t, dtype) -> bool:
    return isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.numel() > 0


def normalize_tensor(t, contrast_variable=0):
    """((t - min)/(max - min) + c)/(1 + c), gradient through min and max (render.py:5-6).

    CUDA float32 tensors with a numeric ``contrast_variable`` use the HIP
    kernels (same FP32 operation order, so the forward values match torch's
    formula); other tensors use the formula in torch."""
    if _hip_ok(t, torch.float32) and isinstance(contrast_variable, (int, float)):
        return _NormalizeHIP.apply(t.contiguous(), float(contrast_variable))
    return ((t - t.min()) / (t.max() - t.min()) + contrast_variable) / (1 + contrast_variable)


def compute_box_size(verts, offset=0.1):
    """Cubic box around ``verts`` (render.py:9-18).

    Per axis the range is extended by ``offset/2`` of its extent on both
    sides; then every axis gets the smallest minimum and largest maximum over
    the three axes.  Returns a float64 numpy array [[x0, x1], [y0, y1], [z0, z1]].
    """
    verts = np.asarray(verts)
    box_size = np.zeros((3, 2))
    extent = verts.max(axis=0) - verts.min(axis=0)
    box_size[:, 1] = verts.max(axis=0) + extent * offset / 2
    box_size[:, 0] = verts.min(axis=0) - extent * offset / 2
    box_size[:, 0] = box_size[:, 0].min(axis=0)
    box_size[:, 1] = box_size[:, 1].max(axis=0)
    return box_size


def spatial_frequency_axes(box_size: torch.Tensor, box_shape, dtype, device):
    """xi_d[k] = (2k - (n_d - 1)) * pi / (box[d,1] - box[d,0]) (render.py:153-172).

    Computed with the reference's operation order in ``dtype`` so the band
    frequencies are bit-identical to the reference's.
    """
    axes = []
    for d, n in enumerate(box_shape):
        s = np.pi / (box_size[d, 1] - box_size[d, 0])
        k = torch.arange(n, dtype=dtype, device=device)
        axes.append((2 * k - (n - 1)) * s)
    return tuple(axes)


class Fourier3dMesh(nn.Module):
    """The meshFT layer: a triangle mesh in, its band-limited Fourier transform out.

    Same constructor as render.py:116 (``OTF`` before ``narrowband_thresh``
    positionally; keywords work as well).  ``OTF`` is mutable: the band
    ``OTF > narrowband_thresh`` is derived from ``self.OTF`` at each forward
    and cached until that tensor changes.
    """

    def __init__(self, box_size, box_shape, device='cpu', dtype=torch.float, OTF=None,
                 narrowband_thresh=0.01, *, cache_band=True):
        super().__init__()
        # cache_band (not in the reference): keep the compacted band between
        # forwards while self.OTF, its storage and version counter, the
        # threshold and the axes stay the same; False recompacts every forward
        self.cache_band = cache_band
        if not isinstance(box_size, torch.Tensor):
            box_size = torch.as_tensor(np.asarray(box_size))
        self.box_size = torch.tensor(box_size.clone().detach().cpu().numpy(), device=device,
                                     dtype=dtype)
        self.box_shape = list(box_shape)
        self.dtype = dtype
        self.device = device
        self.xi0, self.xi1, self.xi2 = self._compute_spatial_frequency_grid()
        if OTF is None:
            self.OTF = torch.ones(self.box_shape, device=device, dtype=dtype)
        else:
            self.OTF = OTF
        self.narrowband_thresh = narrowband_thresh

    def forward(self, Verts, Faces, Faces_coeff):
        """Verts [nv, 3] float32, Faces [nf, 3] int64/int32, Faces_coeff [nf]
        -> complex64 tensor of shape box_shape (zero outside the band)."""
        if torch.device(self.device).type == 'cpu':
            # the reference's CPU device (MeshFTPy, render.py:142-144): torch on
            # the host, see cpu_path.py; GPU modules never take this branch
            return meshft_cpu(Verts - self.box_size[:, 0], Faces, Faces_coeff, self.xi0,
                              self.xi1, self.xi2, self.OTF, self.narrowband_thresh)
        return MeshFTHIP.apply(Verts - self.box_size[:, 0], Faces, Faces_coeff, self.xi0,
                               self.xi1, self.xi2, self.OTF, self.narrowband_thresh,
                               self.cache_band)

    def clear_band_cache(self):
        """Forget the band compacted from ``self.OTF`` (needed after writing the
        OTF through ``.data`` or a raw pointer, which torch does not version)."""
        clear_band_cache(self.OTF)

    def _compute_spatial_frequency_grid(self):
        return spatial_frequency_axes(self.box_size, self.box_shape, self.dtype, self.device)


def render_image_from_ftmesh(ftmesh, OTF, box_shape):
    """|ifftn(ftmesh * OTF)| over the whole box (render.py:177-180).

    The inverse FFT is torch's (hipFFT/rocFFT on ROCm); no fftshift, as in the
    reference (the centred grid's linear phase disappears in the modulus).
    """
    ftexp = ftmesh * OTF
    spec = torch.fft.ifftn(ftexp, list(box_shape))
    if _hip_ok(spec, torch.complex64):
        spec = spec.contiguous()
        img = _AbsHIP.apply(spec)
        # the abs pass also reduced min and max; normalize_tensor reuses them
        # while the image is unchanged (same version)
        img._dmic_minmax = (img._dmic_mm, img._version)
        return img
    return torch.abs(spec)


def compute_spatial_frequency_grid(box_shape, box_size, device='cpu', dtype=torch.float):
    """xi_d[k] = (2k - (n_d - 1)) * pi / box_size[d] for box edge lengths
    ``box_size`` (render.py:321-340; same operation order)."""
    axes = []
    for d, n in enumerate(box_shape):
        s = np.pi / box_size[d]
        k = torch.arange(int(n), dtype=dtype, device=device)
        axes.append((2 * k - (int(n) - 1)) * s)
    return tuple(axes)


def center_verts_in_box(verts, box_size, offset=0.4):
    """Scales ``verts`` (numpy [nv, 3], modified in place as in render.py:343-355)
    to fill (1 - offset) of the box and centres them in it."""
    box = np.array(box_size)
    ext_box = box[:, 1] - box[:, 0]
    verts *= min(ext_box / (verts.max(axis=0) - verts.min(axis=0)))
    verts *= 1 - offset
    ext = verts.max(axis=0) - verts.min(axis=0)
    verts += box[:, 1] - verts.max(axis=0) - 0.5 * (ext_box - ext)
    return verts


def _heron(Verts, Faces):
    P = Verts[Faces]
    L = torch.norm(P - P[:, [2, 0, 1]], dim=2)  # side lengths, as render.py:188-190
    h = L.sum(1) / 2
    return (h * (h - L[:, 0]) * (h - L[:, 1]) * (h - L[:, 2])) ** 0.5, P


def get_internal_triangle_and_mesh_areas(Verts, Faces):
    """(triangle areas by Heron's formula, their sum) (render.py:184-199)."""
    trisar, _ = _heron(Verts, Faces)
    return trisar, trisar.sum()


def get_internal_triangle_and_mesh_areas_and_normals(Verts, Faces):
    """(areas, total area, a x b + b x c + c x a per face) (render.py:202-218)."""
    trisar, P = _heron(Verts, Faces)
    nor = (torch.cross(P[:, 0], P[:, 1], dim=1) + torch.cross(P[:, 1], P[:, 2], dim=1)
           + torch.cross(P[:, 2], P[:, 0], dim=1))
    return trisar, trisar.sum(), nor


def compute_volume_manifold(Verts, Faces):
    """-sum_t a . (b x c) / 6 (render.py:221-225)."""
    C = Verts[Faces]
    return -torch.sum(torch.sum(torch.cross(C[:, 1], C[:, 2], dim=1) * C[:, 0], dim=1)) / 6


def fourier_from_mesh_narrowband_autofluorescence(OTF, narrowband_thresh, regint, xi0, xi1, xi2,
                                                  Verts, Faces):
    """Surface plus uniform-interior ("autofluorescence") transform on the band
    OTF > narrowband_thresh (render.py:228-318):

        ftmesh(xi) = sum_t F_t(xi) (2 A_t / sum A + i (xi . nor2_t) / |xi|^2),
        nor2_t = nor_t regint / regvol,

    with phases xi . v of the given vertices (no box shift).  Computed by the
    HIP meshFT kernels (autofluorescence variant, Hermitian halving on
    symmetric bands) and differentiable w.r.t. Verts; the reference evaluates
    the dense [band, faces] expression with torch.  Differences: areas by
    the cross product instead of Heron's formula, and the interior term at
    xi = 0 is its limit -regint instead of 0/0.  Host tensors take the torch
    path of cpu_path.py (FP64 arithmetic, chunked autograd)."""
    if not Verts.is_cuda:
        # host tensors: the torch path of cpu_path.py (the reference's function
        # is device-agnostic torch)
        return meshft_af_cpu(OTF, narrowband_thresh, regint, xi0, xi1, xi2, Verts, Faces)
    return MeshFTAutofluorescenceHIP.apply(Verts, Faces, regint, xi0, xi1, xi2, OTF,
                                           narrowband_thresh)
