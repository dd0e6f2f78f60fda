"""Mesh fitting to a microscopy image through the meshFT layer.

The reference's fit loop (Examples/Gradient_Descent.ipynb, cell 2, :116-253)
renders the current mesh, compares it with the target image and steps the
vertices with a preconditioned Adam.  The plain fit here is the baseline
the learned proposal is judged against:

    loss(V) = mean((normalize(render(meshFT(V))) - target)^2) + lam * lap(V)
    lap(V)  = mean_v |v - mean_{u in N(v)} u|^2          (uniform Laplacian)

stepped by ``torch.optim.Adam(lr=1e-3)``.  Every piece on the GPU is
deterministic: meshFT by construction, the Laplacian through a padded
neighbour table with an explicit adjoint (no scatter-add atomics).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .psf import generate_gaussian_psf
from .render import Fourier3dMesh, normalize_tensor, render_image_from_ftmesh


# ------------------------------------------------------------------ uniform Laplacian
def neighbour_table(faces: np.ndarray, nv: int):
    """(nbr [nv, K] int64 padded with the vertex itself, deg [nv]) of a triangle mesh."""
    nb = [set() for _ in range(nv)]
    for a, b, c in np.asarray(faces):
        nb[a].update((b, c)); nb[b].update((a, c)); nb[c].update((a, b))
    K = max(1, max(len(s) for s in nb))
    tab = np.tile(np.arange(nv)[:, None], (1, K))
    deg = np.zeros(nv, np.int64)
    for v, s in enumerate(nb):
        s = sorted(s)
        tab[v, :len(s)] = s
        deg[v] = len(s)
    return tab, deg


class _UniformLaplacian(torch.autograd.Function):
    """D = V - W V with W the row-normalised adjacency; backward (I - W)^T G."""

    @staticmethod
    def forward(ctx, V, nbr, mask, invdeg):
        mean = (V[nbr] * mask[..., None]).sum(1) * invdeg[:, None]
        ctx.save_for_backward(nbr, mask, invdeg)
        return V - mean

    @staticmethod
    def backward(ctx, G):
        nbr, mask, invdeg = ctx.saved_tensors
        # (W^T G)_u = sum_{v in N(u)} G_v / deg(v)   (adjacency is symmetric)
        Gs = G * invdeg[:, None]
        wtg = (Gs[nbr] * mask[..., None]).sum(1)
        return G - wtg, None, None, None


class UniformLaplacian:
    def __init__(self, faces: np.ndarray, nv: int, device):
        tab, deg = neighbour_table(faces, nv)
        K = tab.shape[1]
        self.nbr = torch.tensor(tab, device=device)
        self.mask = torch.tensor((np.arange(K)[None, :] < deg[:, None]).astype(np.float32),
                                 device=device)
        self.invdeg = torch.tensor(1.0 / np.maximum(deg, 1), dtype=torch.float32, device=device)

    def delta(self, V):
        return _UniformLaplacian.apply(V, self.nbr, self.mask, self.invdeg)

    def __call__(self, V):
        return (self.delta(V) ** 2).sum(1).mean()


# ------------------------------------------------------------------ problem
@dataclass
class FitProblem:
    """A target image, a starting mesh and the render model (config 4)."""
    mesh_ft: Fourier3dMesh
    otf: torch.Tensor
    faces: torch.Tensor
    coeff: torch.Tensor
    target: torch.Tensor
    init_verts: torch.Tensor
    lap: UniformLaplacian
    lam: float = 0.1
    box_shape: tuple = (128, 128, 128)
    info: dict = field(default_factory=dict)

    def render(self, V):
        return normalize_tensor(render_image_from_ftmesh(self.mesh_ft(V, self.faces, self.coeff),
                                                         self.otf, self.box_shape))

    def loss(self, V):
        img = self.render(V)
        return torch.mean((img - self.target) ** 2) + self.lam * self.lap(V)

    def loss_and_grad(self, V):
        Vg = V.detach().requires_grad_(True)
        L = self.loss(Vg)
        (g,) = torch.autograd.grad(L, Vg)
        return L.detach(), g


def make_problem(workload, device, noise_seed=None, lam=0.1, init_verts=None) -> FitProblem:
    """Target image from ``workload`` (its mesh rendered, then noised as
    Poisson(500 img)/500 + N(0, 0.01^2)); the fit starts from
    ``init_verts`` (default: workload.extra['init_verts'])."""
    from .synthetic import noisy_image
    dev = torch.device(device)
    box = torch.tensor(workload.box_size)
    probe = Fourier3dMesh(box, workload.box_shape, device=dev, dtype=torch.float32)
    otf = generate_gaussian_psf(torch.tensor(workload.sigma, device=dev), probe.xi0, probe.xi1,
                                probe.xi2)
    mesh_ft = Fourier3dMesh(box, workload.box_shape, dev, torch.float32, otf, workload.thresh)
    faces = torch.tensor(workload.faces, device=dev)
    coeff = torch.ones(len(workload.faces), device=dev)
    with torch.no_grad():
        clean = normalize_tensor(render_image_from_ftmesh(
            mesh_ft(torch.tensor(workload.verts, device=dev), faces, coeff), otf,
            workload.box_shape))
    seed = workload.seed + 3 if noise_seed is None else noise_seed
    target = torch.tensor(noisy_image(clean.cpu().numpy(), seed), device=dev)
    iv = workload.extra["init_verts"] if init_verts is None else init_verts
    return FitProblem(mesh_ft=mesh_ft, otf=otf, faces=faces, coeff=coeff, target=target,
                      init_verts=torch.tensor(np.asarray(iv, np.float32), device=dev),
                      lap=UniformLaplacian(workload.faces, len(iv), dev), lam=lam,
                      box_shape=tuple(workload.box_shape),
                      info={"seed": workload.seed, "name": workload.name})


# ------------------------------------------------------------------ plain Adam
@dataclass
class FitResult:
    verts: torch.Tensor
    losses: list
    times: list          # wall seconds since the start, after each step (synchronised)
    accepted: int = 0
    proposals: int = 0
    snapshots: list = field(default_factory=list)
    accepted_at: list = field(default_factory=list)  # learned fit: iterations that took a proposal


def adam_direction(opt: torch.optim.Adam, p: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """The normalised step m_hat / (sqrt(v_hat) + eps) that ``opt`` (torch.optim.Adam,
    one parameter group) would take next for parameter ``p`` with gradient ``g``,
    without the learning rate and without changing the optimiser state."""
    group = opt.param_groups[0]
    b1, b2 = group["betas"]
    eps = group["eps"]
    st = opt.state.get(p, {})
    t = int(st["step"]) + 1 if "step" in st else 1
    m = st["exp_avg"] * b1 + (1 - b1) * g if "exp_avg" in st else (1 - b1) * g
    v = st["exp_avg_sq"] * b2 + (1 - b2) * g * g if "exp_avg_sq" in st else (1 - b2) * g * g
    return (m / (1 - b1 ** t)) / ((v / (1 - b2 ** t)).sqrt() + eps)


def fit_adam(problem: FitProblem, steps: int = 1000, lr: float = 1e-3, snapshot_every: int = 0,
             sync_every: int = 1, stop_loss: float | None = None) -> FitResult:
    """Plain Adam on the vertices (config 4).  ``losses[k]`` is the loss at the
    iterate before step k; the last entry is the loss after the final step.
    Every ``snapshot_every`` steps (vertices, gradient, Adam direction) is kept."""
    V = problem.init_verts.clone().requires_grad_(True)
    opt = torch.optim.Adam([V], lr=lr)
    losses, times, snaps = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        opt.zero_grad(set_to_none=True)
        L = problem.loss(V)
        L.backward()
        if snapshot_every and k % snapshot_every == 0:
            snaps.append((k, V.detach().clone(), V.grad.detach().clone(),
                          adam_direction(opt, V, V.grad.detach())))
        opt.step()
        losses.append(L.detach())
        if sync_every and (k + 1) % sync_every == 0:
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            if stop_loss is not None and float(L) <= stop_loss:
                break
    with torch.no_grad():
        losses.append(problem.loss(V).detach())
    torch.cuda.synchronize()
    times.append(time.perf_counter() - t0)
    return FitResult(verts=V.detach(), losses=[float(x) for x in losses], times=times,
                     snapshots=snaps)


# ------------------------------------------------------------------ large steps
def fit_largesteps(problem: FitProblem, steps: int = 100, lr: float = 1e-2,
                   lambda_: float = 50.0) -> FitResult:
    """The reference's preconditioned fit (Examples/Gradient_Descent.ipynb
    cell 2) on this problem's loss: AdamUniform on the vertices with the
    gradient replaced by (I + lambda L)^-2 grad (deltamic_amd.optim)."""
    from .optim import AdamUniform, LargeStepsPreconditioner
    V = problem.init_verts.clone().requires_grad_(True)
    pre = LargeStepsPreconditioner(problem.faces.cpu().numpy(), V.shape[0], lambda_, V.device)
    opt = AdamUniform([V], lr=lr)
    losses, times = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        L = problem.loss(V)
        L.backward()
        with torch.no_grad():
            V.grad = pre.solve(V.grad)
        opt.step()
        losses.append(float(L))
        times.append(time.perf_counter() - t0)
    with torch.no_grad():
        losses.append(float(problem.loss(V)))
    times.append(time.perf_counter() - t0)
    return FitResult(verts=V.detach(), losses=losses, times=times)
