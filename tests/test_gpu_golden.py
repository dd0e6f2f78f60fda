"""The HIP path against the committed golden vectors of tests/golden/,
directly (no oracle in between): mpmath quadrature of one-face meshes over
every branch of the face formula, and the exact unit-cube surface and
autofluorescence transforms.

Tolerances, relative (face values span 1 down to 1e-8 for phase gaps of
2e4 rad): one-face value within 2e-6 |value|, vertex gradient within
1e-5 max|gradient|; cube rel-L2 1e-5 (forward contract), also over the
frequencies whose phase gaps exceed 1000 rad on their own.
"""
import numpy as np
import pytest
import torch

from tests.test_golden import FACES, cplx, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", FACES, ids=[c["case"] for c in FACES])
def test_one_face_mesh_vs_quadrature(device, case):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    print(f"golden face case: {case['case']}", flush=True)
    V = torch.tensor(case["verts"], dtype=torch.float32, device=device)
    Fc = torch.tensor([[0, 1, 2]], dtype=torch.int64, device=device)
    C = torch.ones(1, dtype=torch.float32, device=device)
    X = torch.tensor([case["xi"]], dtype=torch.float32, device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, X)
    got = complex(ft[0].item())
    want = complex(cplx(case["value"]))
    err = abs(got - want)
    print(f"  value {got:.8g} want {want:.8g} err {err:.2e}")
    assert err <= 2e-6 * abs(want) + 1e-15
    g = 0.6 - 0.8j
    go = torch.tensor([g], dtype=torch.complex64, device=device)
    gv = backward_fourier3d_cpp(V, Fc, C, go, ft, X).double().cpu().numpy()
    xi = np.array(case["xi"], dtype=np.float64)
    gw = np.real(np.conj(g) * cplx(case["dvalue_dphase"]))[:, None] * xi[None, :]
    gerr = np.max(np.abs(gv - gw))
    print(f"  gradient err {gerr:.2e} (max |gradient| {np.max(np.abs(gw)):.3e})")
    assert gerr <= 1e-5 * np.max(np.abs(gw)) + 1e-15


def test_unit_cube_vs_exact(device):
    from deltamic_amd.fourier3d_cpp import forward_fourier3d_cpp
    from oracle.oracle import rel_l2
    c = load("cube_surface.json")
    xi = np.array(c["xi"])
    V = torch.tensor(c["verts"], dtype=torch.float32, device=device)
    Fc = torch.tensor(c["faces"], dtype=torch.int64, device=device)
    C = torch.ones(len(c["faces"]), dtype=torch.float32, device=device)
    X = torch.tensor(xi, dtype=torch.float32, device=device)
    surf, solid = cplx(c["surface"]), cplx(c["solid"])
    got = forward_fourier3d_cpp(V, Fc, C, X).cpu().numpy().astype(np.complex128)
    w = c["wide_from"]
    e_all, e_wide = rel_l2(got, surf), rel_l2(got[w:], surf[w:])
    af = forward_fourier3d_cpp(V, Fc, C, X, regint=0.7).cpu().numpy().astype(np.complex128)
    want_af = surf - 0.7 * solid
    want_af[np.all(xi == 0, axis=1)] = 1.0 - 0.7
    e_af = rel_l2(af, want_af)
    print(f"unit cube: rel-L2 {e_all:.2e} (gaps > 1000 rad: {e_wide:.2e}); "
          f"autofluorescence rel-L2 {e_af:.2e}; max abs err {np.max(np.abs(got - surf)):.2e}")
    assert e_all <= 1e-5 and e_wide <= 1e-5 and e_af <= 1e-5
    assert np.max(np.abs(got - surf)) <= 1e-6
