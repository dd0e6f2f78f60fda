"""deltamic_amd: deltaMic's narrow-band mesh Fourier transform (meshFT) as
hand-written HIP kernels for AMD Instinct MI355X (gfx950).

Public names follow the reference's ``deltamic`` package (``from .psf import *``
and ``from .render import *`` in deltamic/__init__.py:1-2).
"""
from .psf import (Anisotropic_gaussian_psf, Gaussian_psf, Gaussian_psf_logsigma,
                  generate_gaussian_psf)
from .render import (Fourier3dMesh, center_verts_in_box, compute_box_size,
                     compute_spatial_frequency_grid, compute_volume_manifold,
                     fourier_from_mesh_narrowband_autofluorescence,
                     get_internal_triangle_and_mesh_areas,
                     get_internal_triangle_and_mesh_areas_and_normals, normalize_tensor,
                     render_image_from_ftmesh, spatial_frequency_axes)
from .hip_class import MeshFTHIP, MeshFTCUDA

__all__ = [
    "Anisotropic_gaussian_psf",
    "Fourier3dMesh",
    "Gaussian_psf",
    "Gaussian_psf_logsigma",
    "MeshFTCUDA",
    "MeshFTHIP",
    "center_verts_in_box",
    "compute_box_size",
    "compute_spatial_frequency_grid",
    "compute_volume_manifold",
    "fourier_from_mesh_narrowband_autofluorescence",
    "get_internal_triangle_and_mesh_areas",
    "get_internal_triangle_and_mesh_areas_and_normals",
    "generate_gaussian_psf",
    "normalize_tensor",
    "render_image_from_ftmesh",
    "spatial_frequency_axes",
]
