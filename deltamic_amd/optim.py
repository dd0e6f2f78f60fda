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

The preconditioner keeps M sparse (CSR adjacency of the vertex graph,
O(nnz) memory) and applies M^-2 as two Jacobi-preconditioned conjugate-
gradient solves with M in one HIP workgroup (csrc/laplace.hip): the
reference's CholeskySolver(M @ M) semantics at any mesh size.  The dense
variant (method="dense": M^-2 formed once per topology with torch.linalg,
applied as one V x V GEMM) is kept for small meshes and as a check.
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


def vertex_adjacency_csr(faces, nv: int):
    """(row_ptr [nv + 1], col [nnz]) int32: each vertex's neighbours in the
    edge graph of a triangle mesh, sorted, every edge in both rows."""
    f = np.asarray(faces, dtype=np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    e = e[e[:, 0] != e[:, 1]]
    key = np.unique(e[:, 0] * nv + e[:, 1])
    src, dst = key // nv, key % nv
    row = np.zeros(nv + 1, np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row).astype(np.int32), dst.astype(np.int32)


class LargeStepsPreconditioner:
    """x = (I + lambda L)^-2 g, the large-steps gradient of the reference's
    fit loop (Gradient_Descent.ipynb cell 2: CholeskySolver(M @ M)).

    method="sparse" (default): CSR adjacency and two conjugate-gradient
    solves with M per application on the GPU (dmic_laplacian_solve, FP64
    iterates, relative residual ``tol``).  method="dense": M^-2 formed once
    (Cholesky factor of the SPD M in FP64, as deltamic/optimize.py:207-212
    forms the inverse), stored in FP32 and applied as one GEMM; O(V^2)
    memory."""

    def __init__(self, faces, nv: int, lambda_: float, device, method: str = "sparse",
                 tol: float = 1e-9, max_iter: int = 5000):
        self.lambda_ = lambda_
        self.method = method
        self.nv = nv
        self.device = torch.device(device)
        if method == "dense":
            M = compute_matrix(faces, nv, lambda_, device, torch.float64)
            Lc = torch.linalg.cholesky(M)
            Minv = torch.cholesky_inverse(Lc)
            self.P = (Minv @ Minv).to(torch.float32).contiguous()
        elif method == "sparse":
            row, col = vertex_adjacency_csr(faces, nv)
            self.row = torch.tensor(row, device=self.device)
            self.col = torch.tensor(col, device=self.device)
            self.tol, self.max_iter = tol, max_iter
            self.iters = torch.zeros(1, dtype=torch.int32, device=self.device)
            self._ws = None
        else:
            raise ValueError(f"unknown method {method!r} (sparse or dense)")

    @torch.no_grad()
    def solve(self, g: torch.Tensor) -> torch.Tensor:
        if self.method == "dense":
            return self.P @ g
        from . import _lib
        if not g.is_cuda:
            raise RuntimeError("the sparse large-steps solve runs on the GPU; pass a CUDA tensor "
                               "or use method='dense'")
        if g.shape != (self.nv, 3):
            raise ValueError(f"gradient must be [{self.nv}, 3], got {tuple(g.shape)}")
        gf = g.detach().to(torch.float32).contiguous()
        x = torch.empty_like(gf)
        L = _lib.lib()
        if self._ws is None:
            self._ws = _lib.workspace(L.dmic_laplacian_solve_workspace(self.nv), self.device)
        _lib.check(L.dmic_laplacian_solve(self.row.data_ptr(), self.col.data_ptr(), self.nv,
                                          float(self.lambda_), 2, gf.data_ptr(), x.data_ptr(),
                                          float(self.tol), int(self.max_iter),
                                          self.iters.data_ptr(), self._ws.data_ptr(),
                                          self._ws.numel(), _lib.stream_of(x)),
                   "dmic_laplacian_solve")
        return x.to(g.dtype)


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
