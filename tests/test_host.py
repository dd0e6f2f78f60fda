"""CPU-side checks: the C-ABI library loads and exports its header, the host
logic (generators, box, grid, API errors) behaves like the reference's."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import torch

from oracle import oracle as O
from deltamic_amd import synthetic as syn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "deltamic.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmic_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from deltamic_amd import _lib
    L = _lib.lib()
    names = _header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"
    assert L.dmic_version() >= 100
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r" T (dmic_\w+)", nm))
    assert exported == set(names)


def test_workspace_sizes_are_sane():
    from deltamic_amd import _lib
    L = _lib.lib()
    # partial-sum buffers are bounded by the launch split (16384 workgroups x
    # 256 lanes x 16 B forward / 80 B adjoint), not by the band size
    # forward: face records and area sums only (72 B per face, any band)
    assert L.dmic_meshft_forward_workspace(10242, 20480, 815400) < 2 << 20
    assert L.dmic_meshft_forward_workspace(10242, 20480, 10 ** 8) == \
        L.dmic_meshft_forward_workspace(10242, 20480, 1)
    # adjoint: 80 B per face and frequency chunk, the chunk count set by the
    # device (on a host without a GPU, the MI355X figure)
    assert L.dmic_meshft_backward_workspace(10242, 20480, 815400) < 128 << 20
    assert L.dmic_meshft_backward_workspace(50002, 100000, 10 ** 7) < 512 << 20
    assert L.dmic_meshft_forward_workspace(642, 1280, 22000) > 0
    assert L.dmic_band_workspace(200, 200, 200) > 0
    assert L.dmic_meshft_forward_workspace(0, 0, 0) > 0


def test_icosphere_counts_closed_and_outward():
    for s, (nv, nf) in {0: (12, 20), 3: (642, 1280), 4: (2562, 5120), 5: (10242, 20480)}.items():
        v, f = syn.icosphere(s)
        assert v.shape == (nv, 3) and f.shape == (nf, 3)
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
        edges = {}
        for a, b, c in f:
            for e in ((a, b), (b, c), (c, a)):
                edges[e] = edges.get(e, 0) + 1
        # every directed edge once and its reverse once: closed, consistently oriented
        assert all(cnt == 1 for cnt in edges.values())
        assert all((b, a) in edges for (a, b) in edges)
        assert nv - len(edges) // 2 + nf == 2
        p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
        vol = np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6
        assert vol > 0


def test_real_sph_harm_orthonormal():
    u, f = syn.icosphere(5)
    p0, p1, p2 = u[f[:, 0]], u[f[:, 1]], u[f[:, 2]]
    c = (p0 + p1 + p2)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    area = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1)
    area *= 4 * np.pi / area.sum()
    th = np.arccos(np.clip(c[:, 2], -1, 1))
    ph = np.arctan2(c[:, 1], c[:, 0])
    Y = np.stack([syn.real_sph_harm(l, m, th, ph) for l in range(4) for m in range(-l, l + 1)])
    G = (Y * area) @ Y.T
    assert np.allclose(G, np.eye(len(Y)), atol=2e-3)


def test_workload_generators():
    w1 = syn.config1()
    assert w1.verts.shape == (642, 3) and w1.faces.shape == (1280, 3)
    assert w1.verts.dtype == np.float32
    assert np.array_equal(syn.config1().verts, w1.verts)  # seeded
    r = np.linalg.norm(w1.verts - w1.verts.mean(0), axis=1)
    assert abs(r.mean() - 0.3) < 0.01
    w2 = syn.config2()
    assert w2.verts.shape == (10242, 3) and w2.faces.shape == (20480, 3)
    L = w2.box_size[0, 1] - w2.box_size[0, 0]
    assert 2.2 < L < 2.9
    assert np.allclose(w2.box_size[:, 0], w2.box_size[0, 0])  # cubic box
    w3 = syn.config3()
    nd = w3.extra["needle_faces"]
    assert len(nd) == 410
    p0, p1, p2 = (w3.verts[w3.faces[nd, j]].astype(np.float64) for j in range(3))
    shortest = np.min(np.stack([np.linalg.norm(p1 - p0, axis=1), np.linalg.norm(p2 - p1, axis=1),
                                np.linalg.norm(p0 - p2, axis=1)]), axis=0)
    assert shortest.max() < 1e-5
    assert w3.box_shape == (256, 256, 128)
    w4 = syn.config4()
    assert w4.faces.shape == (5120, 3) and w4.extra["init_verts"].shape == (2562, 3)


def test_compute_box_size_semantics():
    from deltamic_amd import compute_box_size
    v = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 4.0]])
    b = compute_box_size(v, offset=0.5)
    # per axis: [min - ext*0.25, max + ext*0.25] -> x [-0.25,1.25], y [-0.5,2.5], z [-1,5]
    # then every axis takes the overall min and max (render.py:16-17)
    assert np.allclose(b, [[-1.0, 5.0]] * 3)
    assert np.allclose(compute_box_size(v), [[-0.2, 4.2]] * 3)  # default offset 0.1


def test_frequency_axes_bit_identical_to_restatement():
    from deltamic_amd import Fourier3dMesh
    w = syn.config3()
    m = Fourier3dMesh(torch.tensor(w.box_size), w.box_shape, device="cpu", dtype=torch.float32)
    ax = O.frequency_axes(m.box_size.numpy(), w.box_shape)
    for a, b in zip((m.xi0, m.xi1, m.xi2), ax):
        assert a.dtype == torch.float32
        assert np.array_equal(a.numpy(), b)
    # no zero bin on even grids, symmetric about zero
    assert np.array_equal(m.xi0.numpy(), -m.xi0.numpy()[::-1])
    assert m.OTF.shape == tuple(w.box_shape) and bool((m.OTF == 1).all())
    assert m.narrowband_thresh == 0.01


def test_native_boundary_checks_like_reference():
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    V = torch.zeros(3, 3)
    F = torch.tensor([[0, 1, 2]])
    C = torch.ones(1)
    xi = torch.zeros(4, 3)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        forward_fourier3d_cpp(V, F, C, xi)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        backward_fourier3d_cpp(V, F, C, torch.zeros(4, dtype=torch.complex64),
                               torch.zeros(4, dtype=torch.complex64), xi)


def test_psf_on_cpu_raises():
    """No CPU path: the OTF generator and the learnable PSF modules refuse CPU
    tensors, as the meshFT module does (their GPU numerics are in
    tests/test_gpu_psf.py)."""
    from deltamic_amd import Anisotropic_gaussian_psf, Gaussian_psf, generate_gaussian_psf
    ax = [torch.linspace(-30, 30, n) for n in (5, 6, 7)]
    S = torch.tensor([[1e-3, 2e-4, 0.0], [2e-4, 2e-3, 1e-4], [0.0, 1e-4, 4e-3]])
    with pytest.raises(NotImplementedError):
        generate_gaussian_psf(S, *ax)
    with pytest.raises(NotImplementedError):
        Gaussian_psf((4, 5, 6), torch.tensor([1.0, 1.0, 1.0]), sigma=10.0)()
    with pytest.raises(NotImplementedError):
        Anisotropic_gaussian_psf((4, 5, 6), [1.0, 1.0, 1.0])()
    m = Gaussian_psf((4, 5, 6), torch.tensor([1.0, 2.0, 3.0]), sigma=10.0)
    assert torch.equal(m.sigma.detach(), 10.0 * torch.eye(3))
    assert m.x2[0] == -1.5 and m.x2[-1] == 1.5 and m.Gridxi1.shape == (4, 5, 6)


def test_psf_rejects_wrong_sized_sigma_before_launch():
    """The kernels read a 3 x 3 sigma (6 parameters for the anisotropic
    PSF); other sizes are refused on the host, with or without gradients."""
    from deltamic_amd import generate_gaussian_psf
    from deltamic_amd.psf import _AnisotropicPSF, _GaussianPSF
    ax = [torch.linspace(-30, 30, n) for n in (5, 6, 7)]
    for sig in (torch.ones(4), torch.ones(2, 2, requires_grad=True), torch.ones(10)):
        with pytest.raises(ValueError):
            generate_gaussian_psf(sig, *ax)
    with pytest.raises(ValueError):
        _GaussianPSF.apply(torch.ones(4, requires_grad=True), *ax)
    with pytest.raises(ValueError):
        _AnisotropicPSF.apply(torch.ones(5, requires_grad=True), *ax)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_uniform_laplacian_adjoint_matches_dense_autograd():
    """The padded-neighbour Laplacian (deterministic adjoint) equals the dense
    (I - W) formulation, value and gradient."""
    from deltamic_amd.fit import UniformLaplacian, neighbour_table
    u, f = syn.icosphere(2)
    V = torch.tensor(u + syn.rng(5).normal(0, 0.05, u.shape), dtype=torch.float64,
                     requires_grad=True)
    lap = UniformLaplacian(f, len(u), "cpu")
    tab, deg = neighbour_table(f, len(u))
    lap.mask = lap.mask.double()
    lap.invdeg = torch.tensor(1.0 / deg, dtype=torch.float64)
    val = lap(V)
    (g,) = torch.autograd.grad(val, V)
    W = np.zeros((len(u), len(u)))
    for v in range(len(u)):
        W[v, tab[v, :deg[v]]] = 1.0 / deg[v]
    M = torch.tensor(np.eye(len(u)) - W)
    V2 = V.detach().clone().requires_grad_(True)
    val2 = ((M @ V2) ** 2).sum(1).mean()
    (g2,) = torch.autograd.grad(val2, V2)
    assert abs(float(val) - float(val2)) <= 1e-12
    assert torch.allclose(g, g2, atol=1e-12)


def test_proposal_features_on_a_sphere():
    """Per-vertex features: centred/scaled position, outward unit normal,
    gradient scaled to unit RMS, its 1-ring mean, Adam direction passed
    through; 1-ring smoothing of a displacement field."""
    from deltamic_amd.fit import UniformLaplacian
    from deltamic_amd.proposal import Features
    u, f = syn.icosphere(3)
    V = torch.tensor(2.0 * u + 0.5, dtype=torch.float32)
    lap = UniformLaplacian(f, len(u), "cpu")
    feats = Features(f, len(u), lap, "cpu")
    g = torch.tensor(syn.rng(3).normal(size=u.shape), dtype=torch.float32)
    adir = torch.tensor(syn.rng(4).normal(size=u.shape), dtype=torch.float32)
    x = feats(V, 7.0 * g, adir)
    assert x.shape == (len(u), 15)
    P, n, gg, gm, ad = x[:, 0:3], x[:, 3:6], x[:, 6:9], x[:, 9:12], x[:, 12:15]
    assert torch.allclose(P.pow(2).sum(1).mean(), torch.tensor(1.0), atol=1e-5)
    assert torch.allclose(n.norm(dim=1), torch.ones(len(u)), atol=1e-5)
    assert (torch.sum(n * torch.tensor(u, dtype=torch.float32), 1) > 0.99).all()  # outward
    assert torch.allclose(gg.pow(2).sum(1).mean(), torch.tensor(1.0), atol=1e-5)
    assert torch.equal(ad, adir)
    i = 17
    nb = lap.nbr[i][lap.mask[i] > 0]
    assert torch.allclose(gm[i], gg[nb].mean(0), atol=1e-6)
    # smoothing: one pass is the mean of a value and its 1-ring mean; constants are fixed
    d1 = feats.smooth(g, passes=1)
    assert torch.allclose(d1[i], 0.5 * (g[i] + g[nb].mean(0)), atol=1e-6)
    c = torch.ones_like(g) * torch.tensor([0.3, -1.0, 2.0])
    assert torch.allclose(feats.smooth(c), c, atol=1e-6)
    assert feats.smooth(g).norm() < 0.8 * g.norm()


def test_adam_direction_matches_torch_step():
    """adam_direction predicts torch.optim.Adam's next step exactly, and
    absorbing a gradient advances the moments as a step would."""
    from deltamic_amd.fit import adam_direction
    from deltamic_amd.proposal import _adam_absorb
    gen = syn.rng(5)
    p = torch.tensor(gen.normal(size=(40, 3)), dtype=torch.float32, requires_grad=True)
    q = p.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([p], lr=1e-2)
    opt_q = torch.optim.Adam([q], lr=1e-2)
    for it in range(4):
        g = torch.tensor(gen.normal(size=(40, 3)), dtype=torch.float32)
        pred = -1e-2 * adam_direction(opt, p, g)
        before = p.detach().clone()
        p.grad = g
        opt.step()
        assert torch.allclose(p.detach() - before, pred, atol=1e-7)
        _adam_absorb(opt_q, q, g)
    st, sq = opt.state[p], opt_q.state[q]
    assert int(st["step"]) == int(sq["step"]) == 4
    assert torch.allclose(st["exp_avg"], sq["exp_avg"]) and torch.allclose(st["exp_avg_sq"], sq["exp_avg_sq"])


def test_c_abi_rejects_bad_arguments_before_any_launch():
    """Argument errors come back as codes with a message, with nothing
    launched (these calls never reach the GPU)."""
    from deltamic_amd import _lib
    L = _lib.lib()
    null = None
    # negative sizes
    rc = L.dmic_meshft_forward(null, -1, null, 1, 0, null, null, 0, null, null, 0, null)
    assert rc == -1 and b"negative" in L.dmic_last_error()
    rc = L.dmic_meshft_af_forward(null, 3, null, 1, 1, null, 0.5, null, 4, null, null, 0, null)
    assert rc == -1 and b"null" in L.dmic_last_error()
    rc = L.dmic_meshft_af_backward(null, 3, null, 1, 1, null, 0.5, null, null, 4, null, null, 0,
                                   null)
    assert rc == -1
    rc = L.dmic_psf_gaussian(null, null, null, null, -2, 3, 4, null, null)
    assert rc == -1 and b"negative" in L.dmic_last_error()
    rc = L.dmic_psf_aniso_grad(null, null, null, null, null, null, 2, 2, 2, null, null, 0, null)
    assert rc == -1
    # the no-allocation forward refuses meshes beyond its static scratch before any HIP call
    rc = L.dmic_meshft_forward_noalloc(null, 3, null, 1, 262145, null, null, 4, null, null)
    assert rc == -1 and b"262144" in L.dmic_last_error()
    rc = L.dmic_meshft_forward_noalloc(null, 3, null, 1, 1, null, null, 4, null, null)
    assert rc == -1 and b"null" in L.dmic_last_error()
    # GEMM grid limits: rows (one per vertex) may be many, output columns at most 65535 tiles
    fake = 256  # never dereferenced: the size guard returns first
    rc = L.dmic_gemm_bf16(fake, 0, 1, 1, fake, 0, 1, 1, fake, 1, 16, 16 * 65536, 16, null, null,
                          0, 0, null)
    assert rc == -1 and b"65535 column tiles" in L.dmic_last_error()
    # the adjoint refuses a mesh whose grid (face blocks x frequency chunks x
    # 256 threads) would not fit in 32 bits, before any launch: 10^8 faces
    rc = L.dmic_meshft_backward(fake, 3, fake, 1, 100_000_000, fake, fake, fake, fake, 1_000_000,
                                fake, fake, 1 << 62, null)
    assert rc == -1 and b"exceeds the adjoint's grid" in L.dmic_last_error()
    # workspaces: the autofluorescence adjoint keeps three more sums per face
    assert L.dmic_meshft_af_backward_workspace(642, 1280, 22000) > \
        L.dmic_meshft_backward_workspace(642, 1280, 22000)
    assert L.dmic_psf_grad_workspace(200, 200, 200) >= (200 * 200 // 16) * 6 * 8
    assert L.dmic_psf_grad_workspace(-1, 2, 2) == 0
