"""Drop-in for the reference's native module ``deltamic.fourier3d_cpp``.

Same two functions, argument order and meaning as the pybind11 module of
deltamic/cpp/fourier3d_cuda.cpp:112-115, calling the HIP C-ABI
(include/deltamic.h) on torch's current stream instead of launching CUDA
kernels on the legacy default stream followed by cudaDeviceSynchronize
(fourier3d_cuda_kernel.cu:448-459,509-536).

Differences from the reference, all on purpose:
  * faces may be int64 (as the reference requires, .cu:450) or int32;
  * an empty band (n = 0) or an empty mesh returns zeros instead of a
    zero-block launch;
  * any vertex valence works (the reference caps it at 30, .cu:487);
  * float64 inputs (the reference's AT_DISPATCH_FLOATING_TYPES double
    case) go to the _f64 entry points: FP64 vertices and frequencies at the
    boundary, complex128 / float64 results, the same FP32 arithmetic with
    FP64 phases inside as for float32 (accuracy of the float32 path).
"""
from __future__ import annotations

import torch

from . import _lib


def _check_input(x: torch.Tensor, name: str) -> None:
    # the AT_ASSERTM checks of fourier3d_cuda.cpp:23-25 (RuntimeError in Python)
    if not x.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not x.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _check_mesh(mesh_vertices, mesh_faces, faces_coeff):
    _check_input(mesh_vertices, "mesh_vertices")
    _check_input(mesh_faces, "mesh_faces")
    _check_input(faces_coeff, "faces_coeff")
    if mesh_vertices.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"meshFT takes float32 or float64 vertices, got {mesh_vertices.dtype}")
    if faces_coeff.dtype != mesh_vertices.dtype:
        raise TypeError(f"faces_coeff must be {mesh_vertices.dtype} like the vertices, got "
                        f"{faces_coeff.dtype}")
    if mesh_faces.dtype not in (torch.int64, torch.int32):
        raise TypeError(f"mesh_faces must be int64 or int32, got {mesh_faces.dtype}")
    if mesh_vertices.dim() != 2 or mesh_vertices.shape[1] != 3:
        raise ValueError(f"mesh_vertices must be [nv, 3], got {tuple(mesh_vertices.shape)}")
    if mesh_faces.dim() != 2 or mesh_faces.shape[1] != 3:
        raise ValueError(f"mesh_faces must be [nf, 3], got {tuple(mesh_faces.shape)}")
    if faces_coeff.shape != (mesh_faces.shape[0],):
        raise ValueError("faces_coeff must have one entry per face")
    dev = mesh_vertices.device
    for t, nm in ((mesh_faces, "mesh_faces"), (faces_coeff, "faces_coeff")):
        if t.device != dev:
            raise RuntimeError(f"{nm} is on {t.device}, mesh_vertices on {dev}")


def validate_faces(mesh_faces: torch.Tensor, nv: int) -> None:
    """Raises IndexError if a face index lies outside [0, nv) (one device sync).

    The kernels never read out of bounds (a bad index reads the origin and
    sets a flag); this makes the error loud, once per faces tensor version.
    """
    key = (mesh_faces._version, nv)
    if getattr(mesh_faces, "_dmic_valid", None) == key:
        return
    if mesh_faces.numel():
        lo, hi = torch.aminmax(mesh_faces)
        lo, hi = int(lo), int(hi)
        if lo < 0 or hi >= nv:
            raise IndexError(f"face index range [{lo}, {hi}] outside [0, {nv})")
    mesh_faces._dmic_valid = key


def forward_fourier3d_cpp(mesh_vertices: torch.Tensor, mesh_faces: torch.Tensor,
                          faces_coeff: torch.Tensor, xi_list: torch.Tensor,
                          out: torch.Tensor | None = None,
                          regint: float | None = None) -> torch.Tensor:
    """ftmesh_list[n] complex64 (fourier3d_cuda.cpp:27-42, .cu:421-462).

    ``out`` (optional, complex64, at least n entries) receives the values in
    its first n entries.  ``regint`` (not in the reference's module) selects
    the autofluorescence variant of render.py:228-318: face weights
    w_t + i (xi . nor2_t) / |xi|^2 with nor2 = nor regint / regvol."""
    _check_mesh(mesh_vertices, mesh_faces, faces_coeff)
    _check_input(xi_list, "xi_list")
    dt = mesh_vertices.dtype
    if xi_list.dtype != dt or xi_list.dim() != 2 or xi_list.shape[1] != 3:
        raise TypeError(f"xi_list must be {dt} [n, 3] like the vertices")
    ct = torch.complex64 if dt == torch.float32 else torch.complex128
    nv, nf, n = mesh_vertices.shape[0], mesh_faces.shape[0], xi_list.shape[0]
    validate_faces(mesh_faces, nv)
    dev = mesh_vertices.device
    if out is None:
        out = torch.empty((n,), dtype=ct, device=dev)
    elif (out.dtype != ct or not out.is_contiguous() or out.numel() < n or out.device != dev):
        raise ValueError(f"out must be a contiguous {ct} tensor with at least n entries")
    L = _lib.lib()
    nbytes = L.dmic_meshft_forward_workspace(nv, nf, n)
    ws = _lib.workspace(nbytes, dev)
    is64 = int(mesh_faces.dtype == torch.int64)
    if dt == torch.float64:
        if regint is not None:
            raise TypeError("the autofluorescence variant takes float32 inputs")
        rc = L.dmic_meshft_forward_f64(
            mesh_vertices.data_ptr(), nv, mesh_faces.data_ptr(), is64, nf, faces_coeff.data_ptr(),
            xi_list.data_ptr(), n, out.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_of(out))
        _lib.check(rc, "dmic_meshft_forward_f64")
    elif regint is None:
        rc = L.dmic_meshft_forward(
            mesh_vertices.data_ptr(), nv, mesh_faces.data_ptr(), is64, nf, faces_coeff.data_ptr(),
            xi_list.data_ptr(), n, out.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_of(out))
        _lib.check(rc, "dmic_meshft_forward")
    else:
        rc = L.dmic_meshft_af_forward(
            mesh_vertices.data_ptr(), nv, mesh_faces.data_ptr(), is64, nf, faces_coeff.data_ptr(),
            float(regint), xi_list.data_ptr(), n, out.data_ptr(), ws.data_ptr(), ws.numel(),
            _lib.stream_of(out))
        _lib.check(rc, "dmic_meshft_af_forward")
    return out


def backward_fourier3d_cpp(mesh_vertices: torch.Tensor, mesh_faces: torch.Tensor,
                           faces_coeff: torch.Tensor, grad_output_list: torch.Tensor,
                           ftmesh_list: torch.Tensor, xi_list: torch.Tensor,
                           regint: float | None = None) -> torch.Tensor:
    """gradVerts[nv, 3] float32 (fourier3d_cuda.cpp:44-65, .cu:466-539); with
    ``regint``, the gradient of the autofluorescence variant."""
    _check_mesh(mesh_vertices, mesh_faces, faces_coeff)
    _check_input(grad_output_list, "grad_output_list")
    _check_input(ftmesh_list, "ftmesh_list")
    _check_input(xi_list, "xi_list")
    dt = mesh_vertices.dtype
    if xi_list.dtype != dt or xi_list.dim() != 2 or xi_list.shape[1] != 3:
        raise TypeError(f"xi_list must be {dt} [n, 3] like the vertices")
    ct = torch.complex64 if dt == torch.float32 else torch.complex128
    n = xi_list.shape[0]
    if grad_output_list.dtype != ct or grad_output_list.shape != (n,):
        raise TypeError(f"grad_output_list must be {ct} [n]")
    nv, nf = mesh_vertices.shape[0], mesh_faces.shape[0]
    validate_faces(mesh_faces, nv)
    dev = mesh_vertices.device
    grad = torch.empty((nv, 3), dtype=dt, device=dev)
    L = _lib.lib()
    is64 = int(mesh_faces.dtype == torch.int64)
    if dt == torch.float64:
        if regint is not None:
            raise TypeError("the autofluorescence variant takes float32 inputs")
        nbytes = L.dmic_meshft_backward_workspace(nv, nf, n)
        ws = _lib.workspace(nbytes, dev)
        rc = L.dmic_meshft_backward_f64(
            mesh_vertices.data_ptr(), nv, mesh_faces.data_ptr(), is64, nf, faces_coeff.data_ptr(),
            grad_output_list.data_ptr(), ftmesh_list.data_ptr(), xi_list.data_ptr(), n,
            grad.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_of(grad))
        _lib.check(rc, "dmic_meshft_backward_f64")
    elif regint is None:
        nbytes = L.dmic_meshft_backward_workspace(nv, nf, n)
        ws = _lib.workspace(nbytes, dev)
        rc = L.dmic_meshft_backward(
            mesh_vertices.data_ptr(), nv, mesh_faces.data_ptr(), is64, nf, faces_coeff.data_ptr(),
            grad_output_list.data_ptr(), ftmesh_list.data_ptr(), xi_list.data_ptr(), n,
            grad.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_of(grad))
        _lib.check(rc, "dmic_meshft_backward")
    else:
        nbytes = L.dmic_meshft_af_backward_workspace(nv, nf, n)
        ws = _lib.workspace(nbytes, dev)
        rc = L.dmic_meshft_af_backward(
            mesh_vertices.data_ptr(), nv, mesh_faces.data_ptr(), is64, nf, faces_coeff.data_ptr(),
            float(regint), grad_output_list.data_ptr(), xi_list.data_ptr(), n, grad.data_ptr(),
            ws.data_ptr(), ws.numel(), _lib.stream_of(grad))
        _lib.check(rc, "dmic_meshft_af_backward")
    return grad
