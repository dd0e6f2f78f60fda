"""The ``device='cpu'`` path of Fourier3dMesh (deltamic_amd/cpu_path.py): the
reference's CPU device (MeshFTPy, deltamic/render.py:36-109) against the
committed golden vectors (mpmath quadrature of single faces, the exact unit
cube) and the FP64 oracle on the small config-1 workload.  Runs on the host;
no GPU involved."""
import json
import os

import numpy as np
import pytest
import torch

from deltamic_amd import Fourier3dMesh
from deltamic_amd import synthetic as syn
from deltamic_amd.cpu_path import MeshFTCPU, band_values
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def cplx(pair):
    a = np.asarray(pair, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


FACES = load("single_faces.json")


@pytest.mark.parametrize("case", FACES, ids=[c["case"] for c in FACES])
def test_one_face_value_and_gradient_vs_quadrature(case):
    V = torch.tensor(case["verts"], dtype=torch.float64, requires_grad=True)
    xi = torch.tensor([case["xi"]], dtype=torch.float64)
    F = torch.tensor([[0, 1, 2]])
    val = MeshFTCPU.apply(V, F, torch.ones(1, dtype=torch.float64), xi)
    want = cplx(case["value"])
    assert abs(val[0].item() - want) <= 1e-10 * max(abs(want), 1e-3)
    g = torch.tensor([0.6 - 0.8j], dtype=torch.complex128)
    val.backward(g)
    # d/dv_j Re(conj(g) value) = Re(conj(g) dvalue/dx_j) xi (one face: constant weight)
    gw = np.real(np.conj(0.6 - 0.8j) * cplx(case["dvalue_dphase"]))[:, None] * np.array([case["xi"]])
    assert np.max(np.abs(V.grad.numpy() - gw)) <= 1e-9 * max(np.max(np.abs(gw)), 1e-3)


def test_unit_cube_vs_exact_transform():
    c = load("cube_surface.json")
    V = torch.tensor(c["verts"], dtype=torch.float64)
    F = torch.tensor(c["faces"])
    xi = torch.tensor(c["xi"], dtype=torch.float64)
    got = band_values(V, F, torch.ones(len(F), dtype=torch.float64), xi).numpy()
    assert np.max(np.abs(got - cplx(c["surface"]))) <= 1e-11


def _config1_module(dtype):
    w = syn.config1()
    m = Fourier3dMesh(torch.tensor(w.box_size), w.box_shape, device="cpu", dtype=dtype)
    otf = torch.tensor(O.gaussian_otf(np.array(w.sigma), m.xi0.numpy(), m.xi1.numpy(),
                                      m.xi2.numpy()), dtype=dtype)
    m.OTF, m.narrowband_thresh = otf, w.thresh
    return w, m


@pytest.mark.parametrize("dtype,ftol,gtol", [(torch.float32, 1e-6, 1e-6),
                                             (torch.float64, 1e-11, 1e-10)])
def test_config1_module_vs_oracle(dtype, ftol, gtol):
    """Forward on the whole band and Verts.grad through torch autograd against
    the FP64 oracle on the same inputs (float32: the complex64 output's
    rounding; float64: FP64 arithmetic)."""
    w, m = _config1_module(dtype)
    V = torch.tensor(w.verts, dtype=dtype, requires_grad=True)
    ft = m(V, torch.tensor(w.faces), torch.ones(len(w.faces), dtype=dtype))
    assert ft.dtype == (torch.complex64 if dtype == torch.float32 else torch.complex128)
    assert tuple(ft.shape) == tuple(w.box_shape)
    mask = m.OTF.abs() > m.narrowband_thresh
    assert bool((ft[~mask] == 0).all())
    g0, g1, g2 = torch.meshgrid(m.xi0, m.xi1, m.xi2, indexing="ij")
    xi = torch.stack((g0[mask], g1[mask], g2[mask]), 1).to(torch.float64).numpy()
    Vs = (V.detach() - m.box_size[:, 0]).to(torch.float64).numpy()
    ref = O.acc_forward(Vs, w.faces, w.coeff, xi, fp32_inputs=(dtype == torch.float32))
    assert O.rel_l2(ft[mask].detach().numpy(), ref) <= ftol
    g = syn.rng(w.seed + 7)
    gb = g.standard_normal(len(xi)) + 1j * g.standard_normal(len(xi))
    G = torch.zeros(ft.shape, dtype=ft.dtype)
    G[mask] = torch.tensor(gb, dtype=ft.dtype)
    ft.backward(G)
    refg = O.acc_backward(Vs, w.faces, w.coeff, xi, gb.astype(np.complex128),
                          fp32_inputs=(dtype == torch.float32))
    assert O.rel_l2(V.grad.to(torch.float64).numpy(), refg) <= gtol


def test_band_is_abs_otf_above_threshold_and_degenerate_faces_stay_finite():
    """|OTF| > thresh selects the band (render.py:49: negative OTF values
    count by magnitude); a zero-area face and faces nearly perpendicular to
    xi (Taylor branch) give finite values and gradients."""
    box = torch.tensor([[0.0, 1.0]] * 3)
    m = Fourier3dMesh(box, (6, 6, 6), device="cpu", dtype=torch.float64)
    otf = torch.zeros(6, 6, 6, dtype=torch.float64)
    otf[1, 2, 3] = -0.5
    otf[4, 4, 4] = 0.5
    otf[0, 0, 0] = 1e-3
    m.OTF, m.narrowband_thresh = otf, 1e-2
    V = torch.tensor([[0.1, 0.1, 0.1], [0.5, 0.1, 0.1], [0.1, 0.6, 0.1], [0.1, 0.1, 0.7],
                      [0.3, 0.3, 0.3]], dtype=torch.float64, requires_grad=True)
    F = torch.tensor([[0, 1, 2], [0, 2, 3], [0, 3, 1], [1, 3, 2], [4, 4, 4]])
    ft = m(V, F, torch.ones(5, dtype=torch.float64))
    nz = ft != 0
    assert bool(nz[1, 2, 3]) and bool(nz[4, 4, 4]) and int(nz.sum()) == 2
    (ft.real.sum() + ft.imag.sum()).backward()
    assert torch.isfinite(V.grad).all()
    # all phases within 1e-3 rad of each other: the Taylor branch
    xi = torch.tensor([[1e-3, -2e-3, 1.5e-3]], dtype=torch.float64)
    Vt = V.detach().requires_grad_(True)
    val = MeshFTCPU.apply(Vt, F[:4], torch.ones(4, dtype=torch.float64), xi)
    ref = O.acc_forward(V.detach().numpy(), F[:4].numpy(), np.ones(4), xi.numpy(),
                        fp32_inputs=False)
    assert O.rel_l2(val.detach().numpy(), ref) <= 1e-12
    val.real.sum().backward()
    assert torch.isfinite(Vt.grad).all()


def test_autofluorescence_unit_cube_vs_exact_transform():
    """Surface minus regint times the solid cube's transform (the interior
    term at xi = 0 is its limit -regint)."""
    c = load("cube_surface.json")
    V = torch.tensor(c["verts"], dtype=torch.float64)
    F = torch.tensor(c["faces"])
    xi = torch.tensor(c["xi"], dtype=torch.float64)
    got = band_values(V, F, torch.ones(len(F), dtype=torch.float64), xi, regint=0.7).numpy()
    want = cplx(c["surface"]) - 0.7 * cplx(c["solid"])
    want[np.all(np.array(c["xi"]) == 0, axis=1)] = 1.0 - 0.7
    assert np.max(np.abs(got - want)) <= 1e-10


def test_autofluorescence_on_host_vs_oracle_and_differences():
    """fourier_from_mesh_narrowband_autofluorescence on host tensors: band
    values against the FP64 oracle (config 1), zeros outside OTF > thresh, and
    Verts.grad from autograd against central differences of the same FP64
    function on seeded band frequencies and vertices."""
    from deltamic_amd import fourier_from_mesh_narrowband_autofluorescence
    w, m = _config1_module(torch.float64)
    V = torch.tensor(w.verts, dtype=torch.float64)
    F = torch.tensor(w.faces)
    box = fourier_from_mesh_narrowband_autofluorescence(m.OTF, w.thresh, 0.5, m.xi0, m.xi1,
                                                        m.xi2, V, F)
    mask = m.OTF > w.thresh
    assert box.dtype == torch.complex128 and bool((box[~mask] == 0).all())
    g0, g1, g2 = torch.meshgrid(m.xi0, m.xi1, m.xi2, indexing="ij")
    xi = torch.stack((g0[mask], g1[mask], g2[mask]), 1)
    ref = O.acc_forward_af(w.verts, w.faces, np.ones(len(w.faces)), 0.5, xi.numpy(),
                           fp32_inputs=False)
    assert O.rel_l2(box[mask].numpy(), ref) <= 1e-11
    r = syn.rng(w.seed + 11)
    sel = torch.tensor(np.sort(r.choice(len(xi), 256, replace=False)))
    xs = xi[sel]
    g = torch.tensor(r.standard_normal(256) + 1j * r.standard_normal(256))
    ones = torch.ones(len(F), dtype=torch.float64)

    def loss(Vx):
        val = MeshFTCPU.apply(Vx, F, ones, xs, 0.5)
        return (val.real * g.real + val.imag * g.imag).sum()

    Vg = V.clone().requires_grad_(True)
    loss(Vg).backward()
    h = 1e-6
    verts = r.choice(len(w.verts), 6, replace=False)
    fd, ad = [], []
    with torch.no_grad():
        for i in verts:
            for d in range(3):
                Vp, Vm = V.clone(), V.clone()
                Vp[i, d] += h
                Vm[i, d] -= h
                fd.append(float(loss(Vp) - loss(Vm)) / (2 * h))
                ad.append(float(Vg.grad[i, d]))
    assert np.linalg.norm(np.array(ad) - fd) <= 1e-6 * np.linalg.norm(fd)
