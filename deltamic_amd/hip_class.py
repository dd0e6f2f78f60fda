"""``torch.autograd.Function`` of the GPU meshFT path.

Mirrors ``MeshFTCUDA`` of deltamic/cuda_class.py:10-40 (same arguments, same
saved state, same gradient return ``(gradVerts, None x 7)``), with the band
work done by HIP kernels:

* band compaction ``Filter > thresh`` (no abs, cuda_class.py:21) in
  row-major voxel order without materialising a meshgrid (cuda_class.py:20-23);
* the band is cached on the ``Filter`` tensor and rebuilt only when that
  tensor object, its storage address, its version counter, the threshold or
  the frequency axes change (the reference rebuilds it on every forward).
  Writes through ``Filter.data`` or a raw pointer bypass the version counter:
  after such an edit call ``clear_band_cache(Filter)`` (or
  ``Fourier3dMesh.clear_band_cache()``), or build the module with
  ``cache_band=False`` to recompact on every forward as the reference does;
* scatter into the full complex box (cuda_class.py:27-28) and the gather of
  ``grad_output`` (cuda_class.py:36);
* Hermitian halving: when the band is exactly centrosymmetric (always the
  case for the reference's grid with an even OTF such as the Gaussian), band
  entry n-1-j is the mirror -xi of entry j and ftmesh(-xi) = conj(ftmesh(xi)),
  so the forward evaluates the first (n+1)/2 frequencies and mirrors the rest,
  and the adjoint runs on the same half with the folded upstream gradient
  g[j] + conj(g[n-1-j]).  A band that is not centrosymmetric on an exactly
  antisymmetric grid (e.g. the reference fit loop's |fftshift(fftn(PSF))|
  OTF on an even box, symmetric under k <-> n-k rather than n-1-k) is
  evaluated through its symmetric closure B u -B: half of the closure is
  evaluated, mirrored, and B's entries are gathered from it (the adjoint
  scatters B's upstream gradient into the closure first).
  ``HERMITIAN_HALVING = False`` turns both off.

Unlike cuda_class.py:8 importing this module does not reseed torch's RNG.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch
from torch.autograd import Function

from . import _lib
from .fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp


HERMITIAN_HALVING = True


@dataclass
class Band:
    idx: torch.Tensor       # int32 [n] linear voxel indices, row-major
    xi_list: torch.Tensor   # float32 [n, 3]
    shape: tuple
    key: tuple
    symmetric: bool = False  # entry n-1-j is exactly -xi of entry j
    closure: "Band | None" = None   # symmetric closure B u -B of a non-centrosymmetric band
    pos: torch.Tensor | None = None  # int32 [n]: position of each entry in the closure
    axes: tuple = ()                 # the frequency axes the band was built from
    _xi64: torch.Tensor | None = None

    def freqs(self, dtype) -> torch.Tensor:
        """Band frequencies [n, 3] in ``dtype``: float32 as compacted, float64
        gathered from the axes in FP64 (exact for float32 axes)."""
        if dtype == torch.float32:
            return self.xi_list
        if self._xi64 is None:
            i = self.idx.long()
            n1, n2 = self.shape[1], self.shape[2]
            a0, a1, a2 = (a.to(device=self.idx.device, dtype=torch.float64) for a in self.axes)
            self._xi64 = torch.stack([a0[i // (n1 * n2)], a1[(i // n2) % n1], a2[i % n2]],
                                     dim=1).contiguous()
        return self._xi64

    @property
    def n(self) -> int:
        return self.idx.numel()

    @property
    def half(self) -> int:
        return (self.n + 1) // 2

    @property
    def size(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]


def _thresh_value(narrowband_thresh) -> float:
    if isinstance(narrowband_thresh, torch.Tensor):
        return float(narrowband_thresh.detach().float().cpu())
    return float(narrowband_thresh)


def compute_band(Filter: torch.Tensor, thresh: float, xi0: torch.Tensor, xi1: torch.Tensor,
                 xi2: torch.Tensor, closure: bool = True) -> Band:
    """Compacts the band ``Filter > thresh`` (HIP kernels; host syncs for the
    count and the symmetry flag).  With ``closure``, a band that is not
    centrosymmetric on antisymmetric axes also gets its symmetric closure."""
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
    symmetric = False
    if n:
        _lib.check(L.dmic_band_fill(f.data_ptr(), n0, n1, n2, float(thresh), axes[0].data_ptr(),
                                    axes[1].data_ptr(), axes[2].data_ptr(), idx.data_ptr(),
                                    xi_list.data_ptr(), ws.data_ptr(), ws.numel(), st),
                   "dmic_band_fill")
        flag = torch.zeros((1,), dtype=torch.int32, device=dev)
        _lib.check(L.dmic_band_symmetry(f.data_ptr(), n0, n1, n2, float(thresh),
                                        axes[0].data_ptr(), axes[1].data_ptr(), axes[2].data_ptr(),
                                        flag.data_ptr(), st), "dmic_band_symmetry")
        flag_v = int(flag.item())
        symmetric = flag_v == 0
        if flag_v == 1 and closure:
            # mask not centrosymmetric on antisymmetric axes: the closure is the
            # band of max(Filter, Filter mirrored), mirror = flip of all three axes
            fsym = torch.maximum(f, torch.flip(f, dims=(0, 1, 2))).contiguous()
            cb = compute_band(fsym, thresh, xi0, xi1, xi2, closure=False)
            if cb.symmetric:
                pos = torch.searchsorted(cb.idx, idx, out_int32=True)
                return Band(idx=idx, xi_list=xi_list, shape=(n0, n1, n2), key=(),
                            symmetric=False, closure=cb, pos=pos, axes=(xi0, xi1, xi2))
    return Band(idx=idx, xi_list=xi_list, shape=(n0, n1, n2), key=(), symmetric=symmetric,
                axes=(xi0, xi1, xi2))


def band_for(Filter: torch.Tensor, narrowband_thresh, xi0, xi1, xi2, cache: bool = True) -> Band:
    """Band of ``Filter``, cached on the tensor object itself (``cache=False``:
    recompacted on every call, the reference's behaviour)."""
    thresh = _thresh_value(narrowband_thresh)
    if not cache:
        return compute_band(Filter, thresh, xi0, xi1, xi2)
    key = (Filter.data_ptr(), Filter._version, thresh, id(xi0), xi0._version, id(xi1),
           xi1._version, id(xi2), xi2._version)
    cached = getattr(Filter, "_dmic_band", None)
    if cached is not None and cached[0] == key and cached[2] is xi0 and cached[3] is xi1 \
            and cached[4] is xi2:
        return cached[1]
    band = compute_band(Filter, thresh, xi0, xi1, xi2)
    band.key = key
    Filter._dmic_band = (key, band, xi0, xi1, xi2)
    return band


def clear_band_cache(Filter: torch.Tensor) -> None:
    """Drops the band cached on ``Filter``; the next forward recompacts it."""
    if hasattr(Filter, "_dmic_band"):
        del Filter._dmic_band


def _scatter(vals: torch.Tensor, idx: torch.Tensor, size: int) -> torch.Tensor:
    """[size] zeros with out[idx[j]] = vals[j] (complex64 by the HIP kernel)."""
    if vals.dtype == torch.complex64:
        out = torch.empty((size,), dtype=torch.complex64, device=vals.device)
        L = _lib.lib()
        _lib.check(L.dmic_band_scatter(vals.data_ptr(), idx.data_ptr(), idx.numel(),
                                       out.data_ptr(), size, _lib.stream_of(out)),
                   "dmic_band_scatter")
        return out
    out = torch.zeros((size,), dtype=vals.dtype, device=vals.device)
    out[idx.long()] = vals
    return out


def _gather(src: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """src[idx] (complex64 by the HIP kernel)."""
    src = src.contiguous()
    if src.dtype == torch.complex64:
        vals = torch.empty((idx.numel(),), dtype=torch.complex64, device=src.device)
        L = _lib.lib()
        _lib.check(L.dmic_band_gather(src.data_ptr(), idx.data_ptr(), idx.numel(),
                                      vals.data_ptr(), _lib.stream_of(vals)), "dmic_band_gather")
        return vals
    return src.reshape(-1)[idx.long()]


def scatter_band(vals: torch.Tensor, band: Band) -> torch.Tensor:
    return _scatter(vals, band.idx, band.size).reshape(band.shape)


def gather_band(box: torch.Tensor, band: Band, dtype=torch.complex64) -> torch.Tensor:
    if box.dtype != dtype:
        box = box.to(dtype)
    return _gather(box, band.idx)


def _mirror_fill(out: torch.Tensor, n: int) -> None:
    """out[n-1-j] = conj(out[j]) for j < n/2."""
    if out.dtype == torch.complex64:
        L = _lib.lib()
        _lib.check(L.dmic_band_mirror_fill(out.data_ptr(), n, _lib.stream_of(out)),
                   "dmic_band_mirror_fill")
        return
    h = n // 2
    out[n - h:n] = torch.conj(out[:h]).flip(0)


def _mirror_fold(g: torch.Tensor, n: int) -> torch.Tensor:
    """g_half[j] = g[j] + conj(g[n-1-j]) for j < n/2 (the middle entry of an odd band copied)."""
    half = (n + 1) // 2
    if g.dtype == torch.complex64:
        g_half = torch.empty((half,), dtype=torch.complex64, device=g.device)
        L = _lib.lib()
        _lib.check(L.dmic_band_mirror_fold(g.data_ptr(), n, g_half.data_ptr(),
                                           _lib.stream_of(g_half)), "dmic_band_mirror_fold")
        return g_half
    h = n // 2
    g_half = g[:half].clone()
    g_half[:h] += torch.conj(g[n - h:n]).flip(0)
    return g_half


def band_forward(V, Fc, C, band: Band, regint=None) -> torch.Tensor:
    """ftmesh_list [n] complex64 on the band (half band + mirror when symmetric,
    gathered from the half-evaluated closure when the band has one; both
    variants are Hermitian: real face weights and xi . nor2 odd in xi)."""
    if HERMITIAN_HALVING and band.closure is not None:
        return _gather(band_forward(V, Fc, C, band.closure, regint=regint), band.pos)
    xi = band.freqs(V.dtype)
    if not (HERMITIAN_HALVING and band.symmetric and band.n > 1):
        return forward_fourier3d_cpp(V, Fc, C, xi, regint=regint)
    ct = torch.complex64 if V.dtype == torch.float32 else torch.complex128
    out = torch.empty((band.n,), dtype=ct, device=V.device)
    forward_fourier3d_cpp(V, Fc, C, xi[:band.half], out=out, regint=regint)
    _mirror_fill(out, band.n)
    return out


def band_backward(V, Fc, C, grad_list, ftmesh_list, band: Band, regint=None) -> torch.Tensor:
    """gradVerts [nv, 3] from the band upstream gradient (folded to the half band when symmetric;
    scattered into the closure first when the band has one)."""
    if HERMITIAN_HALVING and band.closure is not None:
        cb = band.closure
        g_c = _scatter(grad_list, band.pos, cb.n)
        # the adjoint does not read ftmesh_list (it rebuilds the area term from
        # its own per-face sums), so the closure gradient stands in for it
        return band_backward(V, Fc, C, g_c, g_c, cb, regint=regint)
    xi = band.freqs(V.dtype)
    if not (HERMITIAN_HALVING and band.symmetric and band.n > 1):
        return backward_fourier3d_cpp(V, Fc, C, grad_list, ftmesh_list, xi, regint=regint)
    g_half = _mirror_fold(grad_list, band.n)
    return backward_fourier3d_cpp(V, Fc, C, g_half, ftmesh_list[:band.half], xi[:band.half],
                                  regint=regint)


class MeshFTHIP(Function):
    """HIP-based Function for the Fourier transform of a triangle mesh."""

    @staticmethod
    def forward(ctx, Verts, Faces, Faces_coeff, xi0, xi1, xi2, Filter, narrowband_thresh,
                cache_band=True):
        if Verts.dtype not in (torch.float32, torch.float64):
            raise TypeError("Mesh FT (HIP) takes float32 or float64 vertices, got "
                            + str(Verts.dtype))
        band = band_for(Filter, narrowband_thresh, xi0, xi1, xi2, cache=cache_band)
        V = Verts.detach().contiguous()
        Fc = Faces.contiguous()
        C = Faces_coeff.detach().to(V.dtype).contiguous()
        ftmesh_list = band_forward(V, Fc, C, band)
        ctx.save_for_backward(V, Fc, C, ftmesh_list)
        ctx.band = band
        ftmesh = scatter_band(ftmesh_list, band)
        # output dtype follows Filter (cuda_class.py:27)
        out_dt = torch.complex128 if Filter.dtype == torch.float64 else torch.complex64
        return ftmesh if ftmesh.dtype == out_dt else ftmesh.to(out_dt)

    @staticmethod
    def backward(ctx, grad_output):
        V, Fc, C, ftmesh_list = ctx.saved_tensors
        band = ctx.band
        grad_list = gather_band(grad_output, band, ftmesh_list.dtype)
        gradVerts = band_backward(V, Fc, C, grad_list, ftmesh_list, band)
        return gradVerts, None, None, None, None, None, None, None, None


class MeshFTAutofluorescenceHIP(Function):
    """Autofluorescence variant (render.py:228-318) as an autograd Function:
    forward(Verts, Faces, regint, xi0, xi1, xi2, Filter, thresh) -> box."""

    @staticmethod
    def forward(ctx, Verts, Faces, regint, xi0, xi1, xi2, Filter, narrowband_thresh):
        if Verts.dtype != torch.float32:
            raise TypeError("Mesh FT (HIP) is implemented for float32 only, got " + str(Verts.dtype))
        band = band_for(Filter, narrowband_thresh, xi0, xi1, xi2)
        V = Verts.detach().contiguous()
        Fc = Faces.contiguous()
        C = torch.ones(Fc.shape[0], dtype=torch.float32, device=V.device)
        r = float(regint)
        ftmesh_list = band_forward(V, Fc, C, band, regint=r)
        ctx.save_for_backward(V, Fc, C, ftmesh_list)
        ctx.band, ctx.regint = band, r
        return scatter_band(ftmesh_list, band)

    @staticmethod
    def backward(ctx, grad_output):
        V, Fc, C, ftmesh_list = ctx.saved_tensors
        grad_list = gather_band(grad_output, ctx.band)
        gradVerts = band_backward(V, Fc, C, grad_list, ftmesh_list, ctx.band, regint=ctx.regint)
        return gradVerts, None, None, None, None, None, None, None


# name of the reference's Function, for code that imports it directly
MeshFTCUDA = MeshFTHIP
