"""Preconditioned optimisers of the reference's fit loop, on the GPU.

The reference fits meshes with "large steps" preconditioning
(Examples/Gradient_Descent.ipynb cell 2): the vertex gradient g is replaced
by x solving (I + lambda L)^2 x = g, L the combinatorial graph Laplacian of
the mesh (largesteps.geometry.compute_matrix; its dense form is
deltamic/optimize.py:207-212), and the vertices are stepped by AdamUniform
(largesteps.optimize), Adam whose second-moment normaliser is the maximum
over all entries.  VectorAdam (deltamic/optimize.py:215-257) normalises by
the squared norm of each vertex's gradient vector instead.

``largesteps`` (unpinned in the reference's requirements) is not part of the
reference and not installed here; its published algorithm is restated:
compute_matrix(V, F, lambda) = I + lambda (D - A) with A the vertex
adjacency and D its degree matrix, and AdamUniform's step
p -= lr * m_hat / (eps + max(sqrt(v_hat))).

The preconditioner M^-2 is formed once per topology and lambda as a dense
V x V matrix (Cholesky factor and inverse via torch.linalg on the GPU,
rocSOLVER), so each fit step applies it as one GEMM (V x V by V x 3, a
plain library GEMM on rocBLAS, HBM-bound: 4 V^2 bytes).
"""
from __future__ import annotations

import numpy as np
import torch


def combinatorial_laplacian(faces, nv: int, device, dtype=torch.float64) -> torch.Tensor:
    """Dense L = D - A of the vertex graph of a triangle mesh (every edge once)."""
    f = torch.as_tensor(np.asarray(faces), device=device, dtype=torch.int64)
    e = torch.cat([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e = torch.cat([e, e.flip(1)])
    A = torch.zeros((nv, nv), dtype=dtype, device=device)
    A[e[:, 0], e[:, 1]] = 1.0
    return torch.diag(A.sum(1)) - A


def compute_matrix(faces, nv: int, lambda_: float, device, dtype=torch.float64) -> torch.Tensor:
    """M = I + lambda L (largesteps.geometry.compute_matrix, combinatorial
    Laplacian), dense."""
    return torch.eye(nv, dtype=dtype, device=device) + lambda_ * combinatorial_laplacian(
        faces, nv, device, dtype)


class LargeStepsPreconditioner:
    """x = (I + lambda L)^-2 g as one GEMM with a precomputed dense matrix
    (deltamic/optimize.py:207-212 computes the same matrix with
    torch.inverse; here via the Cholesky factor of the SPD M, in FP64, stored
    in FP32)."""

    def __init__(self, faces, nv: int, lambda_: float, device):
        M = compute_matrix(faces, nv, lambda_, device, torch.float64)
        Lc = torch.linalg.cholesky(M)
        Minv = torch.cholesky_inverse(Lc)
        self.P = (Minv @ Minv).to(torch.float32).contiguous()
        self.lambda_ = lambda_

    @torch.no_grad()
    def solve(self, g: torch.Tensor) -> torch.Tensor:
        return self.P @ g


class AdamUniform(torch.optim.Optimizer):
    """Adam with one normaliser for all entries: max over sqrt(v_hat)
    (largesteps.optimize.AdamUniform)."""

    def __init__(self, params, lr=0.1, betas=(0.9, 0.999), eps=1e-8):
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps))

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            lr = group["lr"]
            b1, b2 = group["betas"]
            eps = group["eps"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["g1"] = torch.zeros_like(p)
                    st["g2"] = torch.zeros_like(p)
                st["step"] += 1
                g = p.grad
                st["g1"].mul_(b1).add_(g, alpha=1 - b1)
                st["g2"].mul_(b2).add_(g.square(), alpha=1 - b2)
                m1 = st["g1"] / (1 - b1 ** st["step"])
                m2 = st["g2"] / (1 - b2 ** st["step"])
                p.sub_(m1 / (eps + m2.sqrt().max()), alpha=lr)


class VectorAdam(torch.optim.Optimizer):
    """Adam whose second moment is the squared norm of each row (vertex)
    along ``axis`` (deltamic/optimize.py:215-257)."""

    def __init__(self, params, lr=0.1, betas=(0.9, 0.999), eps=1e-8, axis=-1):
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, axis=axis))

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            lr = group["lr"]
            b1, b2 = group["betas"]
            eps = group["eps"]
            axis = group["axis"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["g1"] = torch.zeros_like(p)
                    st["g2"] = torch.zeros_like(p)
                st["step"] += 1
                g = p.grad
                st["g1"].mul_(b1).add_(g, alpha=1 - b1)
                if axis is not None:
                    sq = g.norm(dim=axis, keepdim=True).square().expand_as(g)
                else:
                    sq = g.square()
                st["g2"].mul_(b2).add_(sq, alpha=1 - b2)
                m1 = st["g1"] / (1 - b1 ** st["step"])
                m2 = st["g2"] / (1 - b2 ** st["step"])
                p.sub_(m1 / (eps + m2.sqrt()), alpha=lr)
