"""GPU test of dmic_meshft_forward_noalloc, the forward without workspace or
allocation: its output against dmic_meshft_forward's on two calls in a row,
the second with a smaller input, through the C-ABI.

Tolerances: bitwise where both entry points run the same plan (the small
mesh: the face-split partials fit the static scratch); rel-L2 <= 1e-6 where
the band is too large for the split partials and the no-allocation call sums
every face in one pass (a different FP64 summation order), and <= 1e-5
against the FP64 oracle on 1024 seeded band frequencies.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import threads
from deltamic_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _inputs(w, device):
    from deltamic_amd import generate_gaussian_psf, spatial_frequency_axes
    from deltamic_amd.hip_class import compute_band
    bs = torch.tensor(w.box_size, dtype=torch.float32, device=device)
    axes = spatial_frequency_axes(bs, w.box_shape, torch.float32, device)
    otf = generate_gaussian_psf(torch.tensor(w.sigma, device=device), *axes)
    band = compute_band(otf, w.thresh, *axes)
    Vs = (torch.tensor(w.verts) - bs.cpu()[:, 0]).contiguous()
    return (Vs.to(device), torch.tensor(w.faces, device=device),
            torch.tensor(w.coeff, device=device), band.xi_list, Vs.numpy())


def _both(L, _lib, V, Fc, C, xi, device):
    nv, nf, n = V.shape[0], Fc.shape[0], xi.shape[0]
    ref = torch.empty(n, dtype=torch.complex64, device=device)
    got = torch.full((n,), float("nan"), dtype=torch.complex64, device=device)
    ws = _lib.workspace(L.dmic_meshft_forward_workspace(nv, nf, n), device)
    st = _lib.stream_of(ref)
    _lib.check(L.dmic_meshft_forward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(),
                                     xi.data_ptr(), n, ref.data_ptr(), ws.data_ptr(), ws.numel(),
                                     st), "dmic_meshft_forward")
    _lib.check(L.dmic_meshft_forward_noalloc(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(),
                                             xi.data_ptr(), n, got.data_ptr(), st),
               "dmic_meshft_forward_noalloc")
    torch.cuda.synchronize()
    return got.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_forward_matches_workspace_forward(device):
    from deltamic_amd import _lib
    L = _lib.lib()
    # call 1: 20480 faces, 1,081,808 band frequencies (too many for split partials)
    w2 = syn.config2()
    print("no-alloc call 1: config 2 (20480 faces, 200^3 band)", flush=True)
    V, Fc, C, xi, Vs = _inputs(w2, device)
    got, ref = _both(L, _lib, V, Fc, C, xi, device)
    assert np.isfinite(got).all()
    err = O.rel_l2(got, ref)
    sel = np.sort(syn.rng(w2.seed + 1).choice(len(ref), 1024, replace=False))
    orc = O.acc_forward(Vs, w2.faces, w2.coeff, xi.cpu().numpy()[sel], threads=threads())
    err_o = O.rel_l2(got[sel], orc)
    print(f"config 2: no-alloc vs workspace forward rel-L2 {err:.2e}, vs oracle {err_o:.2e}")
    assert err <= 1e-6 and err_o <= 1e-5
    # call 2, smaller: 1280 faces, 64^3 band; the split partials fit, same plan, same bits
    w1 = syn.config1()
    print("no-alloc call 2: config 1 (1280 faces, 64^3 band)", flush=True)
    V, Fc, C, xi, Vs = _inputs(w1, device)
    got, ref = _both(L, _lib, V, Fc, C, xi, device)
    print(f"config 1: max |no-alloc - workspace| {np.abs(got - ref).max():.1e}")
    assert np.array_equal(got, ref)
    # sizes beyond the scratch come back as an argument error
    rc = L.dmic_meshft_forward_noalloc(V.data_ptr(), V.shape[0], Fc.data_ptr(), 1, 262145,
                                       C.data_ptr(), xi.data_ptr(), xi.shape[0], V.data_ptr(),
                                       _lib.stream_of(V))
    assert rc == -1 and b"262144" in L.dmic_last_error()
