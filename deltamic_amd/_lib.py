"""ctypes binding of the C-ABI in ``include/deltamic.h``.

torch is imported first on purpose: torch ships its own HIP runtime
(``libamdhip64.so.7``) and the HIP library below must bind to that same copy,
which the dynamic loader does by soname once torch has loaded it.
"""
from __future__ import annotations

import ctypes
import os

import torch

# DELTAMIC_AMD_LIB selects another build of the same C-ABI (e.g. an A/B variant)
LIB_PATH = os.environ.get("DELTAMIC_AMD_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libdeltamic_hip.so")

_i64 = ctypes.c_int64
_int = ctypes.c_int
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p
_f32 = ctypes.c_float

# name -> (argtypes, restype); must match include/deltamic.h
SIGNATURES = {
    "dmic_version": ([], _int),
    "dmic_last_error": ([], ctypes.c_char_p),
    "dmic_meshft_forward_workspace": ([_i64, _i64, _i64], _sz),
    "dmic_meshft_forward": (
        [_vp, _i64, _vp, _int, _i64, _vp, _vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_meshft_forward_f64": (
        [_vp, _i64, _vp, _int, _i64, _vp, _vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_meshft_backward_f64": (
        [_vp, _i64, _vp, _int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_meshft_forward_noalloc": ([_vp, _i64, _vp, _int, _i64, _vp, _vp, _i64, _vp, _vp], _int),
    "dmic_meshft_backward_workspace": ([_i64, _i64, _i64], _sz),
    "dmic_meshft_backward": (
        [_vp, _i64, _vp, _int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_meshft_af_forward": (
        [_vp, _i64, _vp, _int, _i64, _vp, _f32, _vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_meshft_af_backward_workspace": ([_i64, _i64, _i64], _sz),
    "dmic_meshft_af_backward": (
        [_vp, _i64, _vp, _int, _i64, _vp, _f32, _vp, _vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_check_index_flag": ([_vp, _vp], _int),
    "dmic_profile_enable": ([_int], _int),
    "dmic_profile_read": ([ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long),
                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)], _int),
    "dmic_band_workspace": ([_i64, _i64, _i64], _sz),
    "dmic_band_count": ([_vp, _i64, _i64, _i64, _f32, _vp, _vp, _sz, _vp], _int),
    "dmic_band_fill": ([_vp, _i64, _i64, _i64, _f32, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp], _int),
    "dmic_band_scatter": ([_vp, _vp, _i64, _vp, _i64, _vp], _int),
    "dmic_band_gather": ([_vp, _vp, _i64, _vp, _vp], _int),
    "dmic_band_symmetry": ([_vp, _i64, _i64, _i64, _f32, _vp, _vp, _vp, _vp, _vp], _int),
    "dmic_band_mirror_fill": ([_vp, _i64, _vp], _int),
    "dmic_band_mirror_fold": ([_vp, _i64, _vp, _vp], _int),
    "dmic_gaussian_otf": ([_vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp], _int),
    "dmic_psf_gaussian": ([_vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp], _int),
    "dmic_psf_grad_workspace": ([_i64, _i64, _i64], _sz),
    "dmic_psf_gaussian_grad": ([_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp, _sz, _vp],
                               _int),
    "dmic_psf_aniso": ([_vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp], _int),
    "dmic_psf_aniso_grad": ([_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, _vp, _sz,
                             _vp], _int),
    "dmic_laplacian_solve_workspace": ([_i64], _sz),
    "dmic_laplacian_solve": ([_vp, _vp, _i64, _f32, _int, _vp, _vp, _f32, _int, _vp, _vp, _sz,
                              _vp], _int),
    "dmic_render_workspace": ([_i64], _sz),
    "dmic_abs_minmax": ([_vp, _i64, _vp, _vp, _vp, _sz, _vp], _int),
    "dmic_minmax": ([_vp, _i64, _vp, _vp, _sz, _vp], _int),
    "dmic_normalize": ([_vp, _i64, _vp, _f32, _vp, _vp], _int),
    "dmic_normalize_backward": ([_vp, _vp, _i64, _vp, _f32, _vp, _vp, _sz, _vp], _int),
    "dmic_abs_backward": ([_vp, _vp, _i64, _vp, _vp], _int),
    "dmic_gemm_bf16": ([_vp, _int, _i64, _i64, _vp, _int, _i64, _i64, _vp, _i64, _i64, _i64, _i64,
                        _vp, _vp, _i64, _int, _vp], _int),
    "dmic_mlp_weight_count": ([], _i64),
    "dmic_mlp_bias_count": ([], _i64),
    "dmic_mlp_forward": ([_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp], _int),
    "dmic_mlp_backward": ([_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp], _int),
    "dmic_adam_step": ([_vp, _vp, _vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _i64, _vp], _int),
    "dmic_to_bf16": ([_vp, _i64, _vp, _vp], _int),
}

_LIB = None


def lib() -> ctypes.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: the HIP library has not been built "
                "(run `python -m deltamic_amd.build` or __graft_entry__.build()); "
                "deltamic_amd has no CPU fallback.")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = handle
    return _LIB


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().dmic_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode(errors='replace') if msg else ''}")


def stream_of(t: torch.Tensor) -> int:
    """Raw hipStream_t of torch's current stream on t's device."""
    return torch.cuda.current_stream(t.device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    """Scratch memory from torch's caching allocator (the library never allocates)."""
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
