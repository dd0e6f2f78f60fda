"""GPU tests of the public API (Fourier3dMesh -> render -> normalize -> backward)
and of the edge cases the reference mishandles (empty band, int32 faces,
valence > 30, zero-area faces, bad indices, wrong device/layout)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import threads
from deltamic_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _module(w, dev):
    from deltamic_amd import Fourier3dMesh, generate_gaussian_psf
    box = torch.tensor(w.box_size)
    probe = Fourier3dMesh(box, w.box_shape, device=dev, dtype=torch.float32)
    otf = generate_gaussian_psf(torch.tensor(w.sigma, device=dev), probe.xi0, probe.xi1, probe.xi2)
    return Fourier3dMesh(box, w.box_shape, dev, torch.float32, otf, w.thresh), otf


def test_module_forward_render_backward_cfg1(device):
    """Whole pipeline vs the oracle: ftmesh box, image, and Verts.grad of
    loss = sum(normalize(render(ftmesh))) (config 1)."""
    from deltamic_amd import normalize_tensor, render_image_from_ftmesh
    w = syn.config1()
    mesh_ft, otf = _module(w, device)
    V = torch.tensor(w.verts, device=device, requires_grad=True)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.ones(len(w.faces), device=device)
    ftmesh = mesh_ft(V, Fc, C)
    assert ftmesh.dtype == torch.complex64 and tuple(ftmesh.shape) == w.box_shape
    img = normalize_tensor(render_image_from_ftmesh(ftmesh, otf, w.box_shape))
    img.sum().backward()

    band = otf._dmic_band[1]
    idx = band.idx.cpu().numpy().astype(np.int64)
    xi = band.xi_list.cpu().numpy()
    Vs = (torch.tensor(w.verts) - mesh_ft.box_size.cpu()[:, 0]).numpy()
    ref_list = O.acc_forward(Vs, w.faces, w.coeff, xi, threads=threads())
    ref_box = O.scatter(ref_list, idx, w.box_shape)
    got_box = ftmesh.detach().cpu().numpy()
    assert np.count_nonzero(got_box.reshape(-1)[np.setdiff1d(np.arange(got_box.size), idx)]) == 0
    assert O.rel_l2(got_box, ref_box) <= 1e-5

    otf_np = otf.cpu().numpy().astype(np.float64)
    ref_img = O.normalize(O.render(ref_box, otf_np, w.box_shape))
    assert O.rel_l2(img.detach().cpu().numpy(), ref_img) <= 1e-5

    # (a) the meshFT adjoint inside the module: the upstream gradient that
    # autograd delivered to MeshFTHIP.backward (recomputed from the GPU's own
    # ftmesh), pushed through the FP64 oracle adjoint
    ft_leaf = ftmesh.detach().clone().requires_grad_(True)
    normalize_tensor(render_image_from_ftmesh(ft_leaf, otf, w.box_shape)).sum().backward()
    g_gpu = ft_leaf.grad.cpu().numpy().reshape(-1)[idx]
    ref_grad = O.acc_backward(Vs, w.faces, w.coeff, xi, g_gpu, threads=threads())
    err = O.rel_l2(V.grad.cpu().numpy(), ref_grad)
    print(f"module cfg1 Verts.grad rel-L2 vs oracle (same upstream): {err:.3e}")
    assert err <= 1e-5

    # (b) end to end against an FP64 pipeline.  sum(|ifftn|) routes the FP32
    # phase noise of voxels with |z| ~ 0 into the upstream gradient
    # (d|z|/dz = z/|z|), so this comparison is limited by the loss, not by
    # meshFT: tolerance 3e-4.
    ft_cpu = torch.tensor(ref_box, dtype=torch.complex128, requires_grad=True)
    normalize_tensor(render_image_from_ftmesh(ft_cpu, torch.tensor(otf_np), w.box_shape)).sum().backward()
    g_box = ft_cpu.grad.numpy().reshape(-1)[idx]
    ref_grad64 = O.acc_backward(Vs, w.faces, w.coeff, xi, g_box, threads=threads())
    err64 = O.rel_l2(V.grad.cpu().numpy(), ref_grad64)
    print(f"module cfg1 Verts.grad rel-L2 vs FP64 pipeline: {err64:.3e}")
    assert err64 <= 3e-4


def test_band_cache_follows_otf_changes(device):
    from deltamic_amd.hip_class import band_for
    w = syn.config1()
    mesh_ft, otf = _module(w, device)
    b1 = band_for(otf, w.thresh, mesh_ft.xi0, mesh_ft.xi1, mesh_ft.xi2)
    assert band_for(otf, w.thresh, mesh_ft.xi0, mesh_ft.xi1, mesh_ft.xi2) is b1
    b2 = band_for(otf, 1e-3, mesh_ft.xi0, mesh_ft.xi1, mesh_ft.xi2)
    assert b2.idx.numel() < b1.idx.numel()
    otf.mul_(0.5)  # in-place change bumps the version counter
    b3 = band_for(otf, 1e-3, mesh_ft.xi0, mesh_ft.xi1, mesh_ft.xi2)
    assert b3 is not b2 and b3.idx.numel() < b2.idx.numel()


def test_band_rebuilt_after_unversioned_otf_edit(device):
    """A write through OTF.data bypasses torch's version counter: the module
    recompacts after clear_band_cache(), and with cache_band=False on every
    forward (the reference's behaviour)."""
    from deltamic_amd import Fourier3dMesh
    w = syn.config1()
    mesh_ft, otf = _module(w, device)
    V = torch.tensor(w.verts, device=device)
    Fc = torch.tensor(w.faces, device=device)
    C = torch.ones(len(w.faces), device=device)
    nz0 = int((mesh_ft(V, Fc, C) != 0).sum())
    otf.data[: w.box_shape[0] // 2] = 0.0  # unversioned edit: half the box leaves the band
    stale = int((mesh_ft(V, Fc, C) != 0).sum())
    assert stale == nz0  # cached band, as documented
    mesh_ft.clear_band_cache()
    nz1 = int((mesh_ft(V, Fc, C) != 0).sum())
    assert 0 < nz1 < nz0
    ref = (otf > w.thresh).sum().item()
    assert nz1 == ref
    fresh_otf = otf.clone()
    nocache = Fourier3dMesh(torch.tensor(w.box_size), w.box_shape, device, torch.float32,
                            fresh_otf, w.thresh, cache_band=False)
    assert int((nocache(V, Fc, C) != 0).sum()) == nz1
    fresh_otf.data.zero_()
    fresh_otf.data[0, 0, 0] = 1.0
    assert int((nocache(V, Fc, C) != 0).sum()) == 1


def test_empty_band_and_empty_mesh(device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    mesh_ft, otf = _module(w, device)
    mesh_ft.narrowband_thresh = 2.0  # nothing passes: N = 0
    V = torch.tensor(w.verts, device=device, requires_grad=True)
    Fc = torch.tensor(w.faces, device=device)
    out = mesh_ft(V, Fc, torch.ones(len(w.faces), device=device))
    assert out.abs().sum().item() == 0.0
    out.real.sum().backward()
    assert V.grad.abs().sum().item() == 0.0
    # zero faces
    xi = torch.randn(17, 3, device=device)
    e = forward_fourier3d_cpp(V.detach(), Fc[:0].contiguous(), torch.ones(0, device=device), xi)
    assert e.abs().sum().item() == 0.0
    gv = backward_fourier3d_cpp(V.detach(), Fc[:0].contiguous(), torch.ones(0, device=device),
                                torch.ones(17, dtype=torch.complex64, device=device), e, xi)
    assert gv.abs().sum().item() == 0.0


def test_int32_faces_match_int64(device):
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    w = syn.config1()
    V = torch.tensor(w.verts, device=device)
    C = torch.ones(len(w.faces), device=device)
    xi = torch.tensor(syn.rng(3).normal(0, 40, (3000, 3)).astype(np.float32), device=device)
    g = torch.randn(3000, dtype=torch.complex64, device=device)
    F64 = torch.tensor(w.faces, device=device)
    F32 = F64.to(torch.int32)
    a = forward_fourier3d_cpp(V, F64, C, xi)
    b = forward_fourier3d_cpp(V, F32, C, xi)
    assert torch.equal(a, b)
    assert torch.equal(backward_fourier3d_cpp(V, F64, C, g, a, xi),
                       backward_fourier3d_cpp(V, F32, C, g, a, xi))


def test_high_valence_and_zero_area_face(device):
    """A fan with a 40-valent centre (the reference caps valence at 30) plus a
    zero-area face (the reference divides by its zero normal)."""
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    k = 40
    ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
    rim = np.stack([np.cos(ang), np.sin(ang), 0.1 * np.sin(3 * ang)], 1)
    verts = np.concatenate([[[0.0, 0.0, 0.3]], rim, [[0.5, 0.5, 0.5]]]).astype(np.float32) + 1.0
    faces = [[0, 1 + i, 1 + (i + 1) % k] for i in range(k)]
    faces.append([1, 2, 1])  # zero area (repeated vertex)
    faces = np.array(faces, dtype=np.int64)
    coeff = np.ones(len(faces), np.float32)
    xi = syn.rng(9).normal(0, 5, (500, 3)).astype(np.float32)
    g = syn.rng(10).standard_normal(500) + 1j * syn.rng(11).standard_normal(500)
    V = torch.tensor(verts, device=device)
    Fc = torch.tensor(faces, device=device)
    C = torch.tensor(coeff, device=device)
    X = torch.tensor(xi, device=device)
    ft = forward_fourier3d_cpp(V, Fc, C, X)
    gv = backward_fourier3d_cpp(V, Fc, C, torch.tensor(g.astype(np.complex64), device=device), ft, X)
    assert torch.isfinite(gv).all()
    assert O.rel_l2(ft.cpu().numpy(), O.acc_forward(verts, faces, coeff, xi)) <= 1e-5
    ref = O.acc_backward(verts, faces, coeff, xi, g.astype(np.complex64))
    assert O.rel_l2(gv.cpu().numpy(), ref) <= 1e-4


def test_bad_face_index_raises_without_fault(device):
    from deltamic_amd import _lib
    from deltamic_amd.fourier3d_cpp import forward_fourier3d_cpp
    V = torch.rand(4, 3, device=device)
    Fc = torch.tensor([[0, 1, 2], [1, 2, 7]], device=device)
    C = torch.ones(2, device=device)
    xi = torch.randn(10, 3, device=device)
    with pytest.raises(IndexError):
        forward_fourier3d_cpp(V, Fc, C, xi)
    # the raw C-ABI reads the bad vertex as the origin and flags it
    L = _lib.lib()
    out = torch.empty(10, dtype=torch.complex64, device=device)
    ws = _lib.workspace(L.dmic_meshft_forward_workspace(4, 2, 10), device)
    st = _lib.stream_of(out)
    rc = L.dmic_meshft_forward(V.data_ptr(), 4, Fc.data_ptr(), 1, 2, C.data_ptr(), xi.data_ptr(), 10,
                               out.data_ptr(), ws.data_ptr(), ws.numel(), st)
    assert rc == 0
    assert L.dmic_check_index_flag(ws.data_ptr(), st) == -4
    assert torch.isfinite(torch.view_as_real(out)).all()


def test_wrong_layout_and_dtype_raise(device):
    from deltamic_amd.fourier3d_cpp import forward_fourier3d_cpp
    V = torch.rand(3, 3, device=device)
    Fc = torch.tensor([[0, 1, 2]], device=device)
    C = torch.ones(1, device=device)
    xi = torch.randn(3, 8, device=device).t()
    with pytest.raises(RuntimeError, match="contiguous"):
        forward_fourier3d_cpp(V, Fc, C, xi)
    with pytest.raises(TypeError):
        forward_fourier3d_cpp(V.double(), Fc, C, xi.contiguous())


def test_render_chain_hip_matches_torch(device):
    """abs + normalize_tensor on the HIP kernels against torch's own ops on
    the same complex box: forward values equal to 1 ulp, gradient w.r.t. the
    box (through ifftn, abs, min/max and the normalisation) to 1e-6, with
    ties at the minimum and maximum sharing their gradient as in torch."""
    from deltamic_amd import normalize_tensor, render_image_from_ftmesh
    g = torch.Generator(device="cpu").manual_seed(5)
    shape = (48, 40, 36)
    ft = torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g))
    otf = torch.rand(shape, generator=g)
    up = torch.randn(shape, generator=g)
    def torch_chain(box, o, c):
        t = torch.abs(torch.fft.ifftn(box * o, list(shape)))
        return ((t - t.min()) / (t.max() - t.min()) + c) / (1 + c)

    for c in (0, 0.2):
        a = ft.to(device).requires_grad_(True)
        img = normalize_tensor(render_image_from_ftmesh(a, otf.to(device), shape), c)
        (ga,) = torch.autograd.grad(img, a, up.to(device))
        # torch's ops on the same FP32 box, and in FP64 (the FFT's own rounding)
        b = ft.to(device).requires_grad_(True)
        ref = torch_chain(b, otf.to(device), c)
        (gb,) = torch.autograd.grad(ref, b, up.to(device))
        d = ft.to(torch.complex128).to(device).requires_grad_(True)
        ref64 = torch_chain(d, otf.double().to(device), c)
        (gd,) = torch.autograd.grad(ref64, d, up.double().to(device))
        ef = (img - ref).abs().max().item()
        eb = ((ga - gb).abs().norm() / gb.abs().norm()).item()
        eb64 = ((ga.to(torch.complex128) - gd).abs().norm() / gd.abs().norm()).item()
        tb64 = ((gb.to(torch.complex128) - gd).abs().norm() / gd.abs().norm()).item()
        print(f"render chain c={c}: forward max abs diff vs torch {ef:.2e}; box gradient rel-L2 "
              f"vs torch FP32 {eb:.2e}, vs FP64 {eb64:.2e} (torch FP32 vs FP64 {tb64:.2e})")
        assert ef <= 1e-6 and eb <= 1e-6 and eb64 <= 2 * tb64 + 1e-7
    # ties: the minimum (0) and maximum (1) each attained several times
    t = torch.tensor([0.0, 1.0, 0.5, 0.0, 1.0, 0.25, 0.0], device=device, requires_grad=True)
    u = torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0], device=device)
    (gh,) = torch.autograd.grad(normalize_tensor(t, 0.3), t, u)
    tc = t.detach().cpu().requires_grad_(True)
    (gr,) = torch.autograd.grad(((tc - tc.min()) / (tc.max() - tc.min()) + 0.3) / 1.3, tc, u.cpu())
    assert torch.allclose(gh.cpu(), gr, rtol=1e-6, atol=1e-7)


def test_cpu_and_gpu_devices_agree_cfg1(device):
    """Fourier3dMesh on device='cpu' (torch on the host, FP64) and on the GPU
    (HIP kernels) with the same inputs and OTF: the complex box and Verts.grad
    for the same upstream gradient agree to the GPU path's accuracy."""
    from deltamic_amd import Fourier3dMesh
    w = syn.config1()
    mesh_gpu, otf = _module(w, device)
    mesh_cpu = Fourier3dMesh(torch.tensor(w.box_size), w.box_shape, "cpu", torch.float32,
                             otf.cpu(), w.thresh)
    Vg = torch.tensor(w.verts, device=device, requires_grad=True)
    Vc = torch.tensor(w.verts, requires_grad=True)
    Fc = torch.tensor(w.faces)
    C = torch.ones(len(w.faces))
    box_g = mesh_gpu(Vg, Fc.to(device), C.to(device))
    box_c = mesh_cpu(Vc, Fc, C)
    assert box_c.dtype == box_g.dtype and box_c.shape == box_g.shape
    ferr = O.rel_l2(box_g.detach().cpu().numpy(), box_c.detach().numpy())
    up = torch.tensor(syn.rng(w.seed + 31).standard_normal(box_c.shape).astype(np.float32))
    (box_g.real * up.to(device)).sum().backward()
    (box_c.real * up).sum().backward()
    gerr = O.rel_l2(Vg.grad.cpu().numpy(), Vc.grad.numpy())
    print(f"cpu vs gpu device, config 1: box rel-L2 {ferr:.2e}, Verts.grad rel-L2 {gerr:.2e}")
    assert ferr <= 1e-5 and gerr <= 1e-4
