"""GPU numerics of the proposal network kernels (csrc/mlp.hip) against a
plain PyTorch FP32 reference of the same ops computed on the CPU (inputs
rounded to bf16 as the matrix cores see them).  Tolerances: 1e-4 relative
for single GEMMs (only the FP32 summation order differs), 1e-2 / 2e-2 for the
whole network and its gradients (activations are rounded to bf16 between
layers on the GPU path; the reference rounds them the same way, so the
remaining difference is FP32 summation order amplified through ReLU)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from deltamic_amd.proposal import N_IN  # noqa: E402


def bf16r(x):
    return x.to(torch.bfloat16).to(torch.float32)


def _gemm(A, B, a_bf16=False, b_bf16=False, transA=False, transB=False, bias=None, mask=None,
          epi=0):
    from deltamic_amd import _lib
    L = _lib.lib()
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    N = B.shape[0] if transB else B.shape[1]
    sa0, sa1 = (1, A.shape[1]) if transA else (A.shape[1], 1)
    sb0, sb1 = (1, B.shape[1]) if transB else (B.shape[1], 1)
    C = torch.empty((M, N), device=A.device)
    rc = L.dmic_gemm_bf16(A.data_ptr(), int(a_bf16), sa0, sa1, B.data_ptr(), int(b_bf16), sb0, sb1,
                          C.data_ptr(), N, M, N, K, _lib.ptr(bias), _lib.ptr(mask),
                          N if mask is not None else 0, epi, _lib.stream_of(C))
    _lib.check(rc, "gemm")
    return C


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (37, 21, 5), (130, 128, 128), (4096, 3, 128),
                                   (128, 128, 3000)])
def test_gemm_mfma_layout_and_values(device, M, N, K):
    g = torch.Generator(device="cpu").manual_seed(M * 7 + N)
    A = torch.randn(M, K, generator=g)
    B = torch.randn(K, N, generator=g)
    ref = bf16r(A) @ bf16r(B)  # CPU reference
    A, B = A.to(device), B.to(device)
    got = _gemm(A, B).cpu()
    assert torch.allclose(got, ref, rtol=1e-4, atol=1e-4 * K ** 0.5)
    # transposed operands through strides
    got_t = _gemm(A.t().contiguous(), B.t().contiguous(), transA=True, transB=True).cpu()
    assert torch.allclose(got_t, ref, rtol=1e-4, atol=1e-4 * K ** 0.5)


def test_gemm_bf16_storage_and_epilogues(device):
    g = torch.Generator(device="cpu").manual_seed(5)
    A = torch.randn(50, 40, generator=g).to(device)
    B = torch.randn(40, 33, generator=g).to(device)
    bias = torch.randn(33, generator=g).to(device)
    Bb = B.to(torch.bfloat16).view(torch.int16).contiguous()
    ref = bf16r(A) @ bf16r(B)
    got = _gemm(A, Bb, b_bf16=True, bias=bias, epi=2)
    assert torch.allclose(got, torch.relu(ref + bias), rtol=1e-4, atol=1e-3)
    mask = torch.randn(50, 33, generator=g).to(device)
    got = _gemm(A, B, mask=mask, epi=3)
    assert torch.allclose(got, torch.where(mask > 0, ref, torch.zeros_like(ref)), rtol=1e-4,
                          atol=1e-3)


def _torch_mlp(net, x):
    """FP32 PyTorch reference on the CPU with bf16-rounded weights and layer inputs."""
    w = bf16r(net.w.cpu())
    b = net.b.cpu()
    x = x.cpu()
    sizes = [(N_IN, 128), (128, 128), (128, 128), (128, 3)]
    Ws, off = [], 0
    for i, o in sizes:
        Ws.append(w[off:off + i * o].view(i, o))
        off += i * o
    bs = [b[0:128], b[128:256], b[256:384], b[384:387]]
    h = x
    for li in range(4):
        h = bf16r(h) @ Ws[li] + bs[li]
        if li < 3:
            h = torch.relu(h)
    return h, Ws, bs


# rows: one column-sum chunk (<= 256 rows), several chunks, and split-K
# weight gradients (>= 1024 rows, 256-row chunks, ragged last chunk)
@pytest.mark.parametrize("rows", [1, 200, 1000, 5000])
def test_mlp_forward_backward_vs_torch(device, rows):
    from deltamic_amd.proposal import ProposalMLP
    net = ProposalMLP(device, seed=3)
    net.b.normal_(0, 0.1)
    g = torch.Generator(device="cpu").manual_seed(9)
    x = torch.randn(rows, N_IN, generator=g).to(device)
    y, acts = net.forward(x, keep_acts=True)
    ref, Ws, bs = _torch_mlp(net, x)
    y = y.cpu()
    assert torch.allclose(y, ref, rtol=1e-2, atol=1e-2 * ref.abs().max().item())
    dy = torch.randn(rows, 3, generator=g).to(device)
    gw, gb = net.backward(x, acts, dy)
    gw, gb, x, dy = gw.cpu(), gb.cpu(), x.cpu(), dy.cpu()
    # autograd of the reference wrt the (bf16-rounded) weights and biases
    Wr = [W.clone().requires_grad_(True) for W in Ws]
    br = [bb.clone().requires_grad_(True) for bb in bs]
    h = x
    for li in range(4):
        h = bf16r(h) @ Wr[li] + br[li]
        if li < 3:
            h = torch.relu(h)
    (h * dy).sum().backward()
    ref_gw = torch.cat([W.grad.reshape(-1) for W in Wr])
    ref_gb = torch.cat([bb.grad for bb in br])
    assert (gw - ref_gw).norm() / ref_gw.norm() < 2e-2
    assert (gb - ref_gb).norm() / ref_gb.norm() < 2e-2
    # the output-layer bias gradient is the plain column sum of dy (FP64 sums)
    assert torch.allclose(gb[-3:].double(), dy.double().sum(0), rtol=1e-6, atol=1e-6)
    # and is reproducible call to call
    gw2, gb2 = net.backward(x.to(device), acts, dy.to(device))
    assert torch.equal(gw2.cpu(), gw) and torch.equal(gb2.cpu(), gb)


def test_mlp_training_reduces_loss_and_is_deterministic(device):
    from deltamic_amd.proposal import ProposalMLP
    g = torch.Generator(device="cpu").manual_seed(1)
    x = torch.randn(4096, N_IN, generator=g).to(device)
    t = 0.02 * torch.tanh(x[:, :3] - 0.5 * x[:, 6:9])
    losses = []
    nets = []
    for rep in range(2):
        net = ProposalMLP(device, seed=11)
        ls = [float(net.train_step(x, t, lr=3e-3)) for _ in range(200)]
        losses.append(ls)
        nets.append(net)
    assert losses[0][-1] < 0.2 * losses[0][0]
    assert losses[0] == losses[1]
    assert torch.equal(nets[0].w, nets[1].w)


def test_adam_matches_torch(device):
    from deltamic_amd import _lib
    L = _lib.lib()
    g = torch.Generator(device="cpu").manual_seed(2)
    w0 = torch.randn(1000, generator=g).to(device)
    w = w0.clone()
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    p = w0.clone().requires_grad_(True)
    opt = torch.optim.Adam([p], lr=1e-2)
    for step in range(1, 6):
        grad = torch.randn(1000, generator=g).to(device)
        _lib.check(L.dmic_adam_step(w.data_ptr(), None, grad.data_ptr(), m.data_ptr(), v.data_ptr(),
                                    1000, 1e-2, 0.9, 0.999, 1e-8, step, _lib.stream_of(w)), "adam")
        p.grad = grad
        opt.step()
    assert torch.allclose(w, p.detach(), rtol=1e-5, atol=1e-6)
