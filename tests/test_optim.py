"""Optimisers of the reference's fit loop (deltamic_amd/optim.py) on the CPU:
AdamUniform and VectorAdam against numpy restatements of their update rules,
the large-steps matrix against its definition, and the preconditioner as a
solver of (I + lambda L)^2 x = g."""
import numpy as np
import torch

from deltamic_amd import synthetic as syn
from deltamic_amd.optim import (AdamUniform, LargeStepsPreconditioner, VectorAdam,
                                combinatorial_laplacian, compute_matrix)


def _run(opt_cls, grads, **kw):
    p = torch.tensor(np.linspace(-1, 1, 30).reshape(10, 3), dtype=torch.float64, requires_grad=True)
    opt = opt_cls([p], **kw)
    for g in grads:
        p.grad = torch.tensor(g)
        opt.step()
    return p.detach().numpy()


def test_adam_uniform_and_vector_adam_updates():
    gen = syn.rng(21)
    grads = [gen.normal(size=(10, 3)) for _ in range(5)]
    lr, b1, b2, eps = 0.05, 0.9, 0.999, 1e-8
    p0 = np.linspace(-1, 1, 30).reshape(10, 3)
    # AdamUniform: one normaliser, the max of sqrt(v_hat) over all entries
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t, g in enumerate(grads, 1):
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (eps + np.sqrt(v / (1 - b2 ** t)).max())
    assert np.allclose(_run(AdamUniform, grads, lr=lr), p, rtol=1e-12, atol=1e-14)
    # VectorAdam: per-row normaliser |g_row|^2
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t, g in enumerate(grads, 1):
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * np.repeat((g * g).sum(1, keepdims=True), 3, axis=1)
        p = p - lr * (m / (1 - b1 ** t)) / (eps + np.sqrt(v / (1 - b2 ** t)))
    assert np.allclose(_run(VectorAdam, grads, lr=lr), p, rtol=1e-12, atol=1e-14)


def test_large_steps_matrix_and_preconditioner():
    u, f = syn.icosphere(2)
    nv = len(u)
    L = combinatorial_laplacian(f, nv, "cpu").numpy()
    assert np.allclose(L, L.T) and np.allclose(L.sum(1), 0.0)
    deg = np.diag(L)
    assert set(np.unique(deg)) == {5.0, 6.0}  # icosphere valences
    M = compute_matrix(f, nv, 50.0, "cpu").numpy()
    assert np.allclose(M, np.eye(nv) + 50.0 * L)
    pre = LargeStepsPreconditioner(f, nv, 50.0, "cpu", method="dense")
    g = torch.tensor(syn.rng(22).normal(size=(nv, 3)), dtype=torch.float32)
    x = pre.solve(g).double().numpy()
    # FP32 storage of the dense inverse: compare with the FP64 solve of M^2 x = g
    x64 = np.linalg.solve(M @ M, g.double().numpy())
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-5
    # constants are fixed points (L annihilates them)
    c = torch.ones((nv, 3))
    assert torch.allclose(pre.solve(c), c, atol=1e-5)


def test_vertex_adjacency_csr_matches_laplacian():
    from deltamic_amd.optim import vertex_adjacency_csr
    u, f = syn.icosphere(2)
    nv = len(u)
    row, col = vertex_adjacency_csr(f, nv)
    L = combinatorial_laplacian(f, nv, "cpu").numpy()
    A = np.zeros((nv, nv))
    for i in range(nv):
        nb = col[row[i]:row[i + 1]]
        assert np.all(np.diff(nb) > 0) and i not in nb
        A[i, nb] = 1.0
    assert np.array_equal(np.diag(L), A.sum(1)) and np.array_equal(L, np.diag(A.sum(1)) - A)
