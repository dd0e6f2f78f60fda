"""Learned per-vertex displacement proposals for the mesh fit.

No counterpart in the reference.  A small MLP (15 -> 128 -> 128 -> 128 -> 3,
bf16 weights, FP32 accumulation; HIP kernels in csrc/mlp.hip) maps
per-vertex features

    position (centred, scaled by the mesh RMS radius), unit vertex normal,
    loss gradient and its 1-ring mean (both multiplied by nv * grad_scale:
    per-vertex gradients scale like the vertex area ~ 1/nv, and grad_scale
    is calibrated on the training trajectories, so the features keep the
    absolute gradient size the step length depends on), and the normalised
    step m_hat / (sqrt(v_hat) + eps) that Adam would take next

to a displacement, which is smoothed over the 1-ring before use.  The
network is trained on the method's own fitting trajectories in two ways:
by regression on the displacement plain Adam makes over the next
``horizon`` steps (train_step), and through the meshFT layer itself
(direct_step): the proposal is applied to a trajectory state and the
network descends log L(V + d), with dL/dd from the HIP meshFT adjoint.  In
the fit (fit_learned) each proposal competes with the plain Adam step it
would replace: it is accepted only if its meshFT + OTF loss is below both
the current loss and the Adam step's loss, otherwise the Adam step stands.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .fit import FitProblem, FitResult, adam_direction

N_IN, N_HID, N_OUT = 15, 128, 3


def vertex_face_table(faces: np.ndarray, nv: int):
    """Incident faces per vertex, padded ([nv, K] int64, mask [nv, K] float32)."""
    inc = [[] for _ in range(nv)]
    for t, tri in enumerate(np.asarray(faces)):
        for v in tri:
            inc[v].append(t)
    K = max(1, max(len(x) for x in inc))
    tab = np.zeros((nv, K), np.int64)
    mask = np.zeros((nv, K), np.float32)
    for v, lst in enumerate(inc):
        tab[v, :len(lst)] = lst
        mask[v, :len(lst)] = 1.0
    return tab, mask


class Features:
    """Per-vertex feature builder for one mesh topology (all torch gathers: deterministic)."""

    def __init__(self, faces_np: np.ndarray, nv: int, lap, device, grad_scale: float | None = None):
        tab, mask = vertex_face_table(faces_np, nv)
        self.faces = torch.tensor(np.asarray(faces_np), device=device)
        self.vf = torch.tensor(tab, device=device)
        self.vf_mask = torch.tensor(mask, device=device)
        self.lap = lap
        self.nv = nv
        # None: scale gradients to unit RMS over the mesh (direction only)
        self.grad_scale = grad_scale

    @torch.no_grad()
    def __call__(self, V: torch.Tensor, grad: torch.Tensor, adir: torch.Tensor) -> torch.Tensor:
        c = V.mean(0, keepdim=True)
        P = V - c
        rad = P.pow(2).sum(1).mean().sqrt().clamp_min(1e-12)
        a, b, cc = V[self.faces[:, 0]], V[self.faces[:, 1]], V[self.faces[:, 2]]
        fn = torch.cross(b - a, cc - a, dim=1)  # area-weighted face normals
        vn = (fn[self.vf] * self.vf_mask[..., None]).sum(1)
        vn = vn / vn.norm(dim=1, keepdim=True).clamp_min(1e-20)
        if self.grad_scale is None:
            g = grad / grad.pow(2).sum(1).mean().sqrt().clamp_min(1e-30)
        else:
            g = grad * (self.nv * self.grad_scale)
        lap = self.lap
        gm = (g[lap.nbr] * lap.mask[..., None]).sum(1) * lap.invdeg[:, None]
        return torch.cat([P / rad, vn, g, gm, adir], dim=1).contiguous()

    @torch.no_grad()
    def smooth(self, d: torch.Tensor, passes: int = 2) -> torch.Tensor:
        """Each pass replaces a displacement by the mean of itself and its 1-ring mean."""
        lap = self.lap
        for _ in range(passes):
            d = 0.5 * (d + (d[lap.nbr] * lap.mask[..., None]).sum(1) * lap.invdeg[:, None])
        return d

    @torch.no_grad()
    def smooth_adjoint(self, g: torch.Tensor, passes: int = 2) -> torch.Tensor:
        """Transpose of smooth: with W the row-normalised adjacency, a pass is
        (I + W)/2 and its transpose (I + W^T)/2, (W^T g)_u = sum_{v in N(u)} g_v / deg v."""
        lap = self.lap
        for _ in range(passes):
            gs = g * lap.invdeg[:, None]
            g = 0.5 * (g + (gs[lap.nbr] * lap.mask[..., None]).sum(1))
        return g


class ProposalMLP:
    """The proposal network: fp32 master parameters, bf16 weight copy, Adam state."""

    def __init__(self, device, seed: int = 0, out_scale: float = 0.02,
                 grad_scale: float | None = None):
        L = _lib.lib()
        self.nw = int(L.dmic_mlp_weight_count())
        self.nb = int(L.dmic_mlp_bias_count())
        self.device = torch.device(device)
        self.out_scale = out_scale
        self.grad_scale = grad_scale  # feature scaling the network was trained with
        g = torch.Generator().manual_seed(seed)
        chunks = []
        for fan_in, fan_out in ((N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)):
            std = math.sqrt(2.0 / fan_in) if fan_out != N_OUT else 0.1 / math.sqrt(fan_in)
            chunks.append(torch.randn(fan_in * fan_out, generator=g) * std)
        self.w = torch.cat(chunks).to(self.device).contiguous()
        self.b = torch.zeros(self.nb, device=self.device)
        self.wb = torch.empty(self.nw, dtype=torch.int16, device=self.device)
        self._sync_bf16()
        self.mw = torch.zeros_like(self.w)
        self.vw = torch.zeros_like(self.w)
        self.mb = torch.zeros_like(self.b)
        self.vb = torch.zeros_like(self.b)
        self.step_count = 0

    def _sync_bf16(self):
        L = _lib.lib()
        _lib.check(L.dmic_to_bf16(self.w.data_ptr(), self.nw, self.wb.data_ptr(),
                                  _lib.stream_of(self.w)), "dmic_to_bf16")

    def state_dict(self):
        return {"w": self.w.cpu(), "b": self.b.cpu(), "out_scale": self.out_scale,
                "grad_scale": -1.0 if self.grad_scale is None else float(self.grad_scale),
                "step": self.step_count}

    def load_state_dict(self, sd):
        if sd["w"].numel() != self.nw or sd["b"].numel() != self.nb:
            raise ValueError(f"proposal weights have {sd['w'].numel()} + {sd['b'].numel()} entries; "
                             f"this network needs {self.nw} + {self.nb}")
        self.w = sd["w"].to(self.device).contiguous()
        self.b = sd["b"].to(self.device).contiguous()
        self.out_scale = float(sd["out_scale"])
        gsc = float(sd.get("grad_scale", -1.0))
        self.grad_scale = None if gsc < 0 else gsc
        self.step_count = int(sd.get("step", 0))
        self._sync_bf16()

    def forward(self, x: torch.Tensor, keep_acts: bool = False):
        """Raw network output [rows, 3] (and hidden activations if keep_acts)."""
        x = x.contiguous()
        rows = x.shape[0]
        y = torch.empty((rows, N_OUT), device=self.device)
        acts = torch.empty((3, rows, N_HID), device=self.device) if keep_acts else None
        scratch = None if keep_acts else torch.empty((2, rows, N_HID), device=self.device)
        L = _lib.lib()
        _lib.check(L.dmic_mlp_forward(self.wb.data_ptr(), self.b.data_ptr(), x.data_ptr(), rows,
                                      y.data_ptr(), _lib.ptr(acts), _lib.ptr(scratch),
                                      _lib.stream_of(y)), "dmic_mlp_forward")
        return (y, acts) if keep_acts else y

    def propose(self, x: torch.Tensor) -> torch.Tensor:
        return self.out_scale * self.forward(x)

    def backward(self, x, acts, dy):
        rows = x.shape[0]
        gw = torch.empty_like(self.w)
        gb = torch.empty_like(self.b)
        scratch = torch.empty((2, rows, N_HID), device=self.device)
        L = _lib.lib()
        _lib.check(L.dmic_mlp_backward(self.wb.data_ptr(), x.data_ptr(), acts.data_ptr(),
                                       dy.contiguous().data_ptr(), rows, gw.data_ptr(),
                                       gb.data_ptr(), scratch.data_ptr(), _lib.stream_of(gw)),
                   "dmic_mlp_backward")
        return gw, gb

    def adam(self, gw, gb, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.step_count += 1
        L = _lib.lib()
        st = _lib.stream_of(self.w)
        _lib.check(L.dmic_adam_step(self.w.data_ptr(), self.wb.data_ptr(), gw.data_ptr(),
                                    self.mw.data_ptr(), self.vw.data_ptr(), self.nw, lr, betas[0],
                                    betas[1], eps, self.step_count, st), "dmic_adam_step")
        _lib.check(L.dmic_adam_step(self.b.data_ptr(), None, gb.data_ptr(), self.mb.data_ptr(),
                                    self.vb.data_ptr(), self.nb, lr, betas[0], betas[1], eps,
                                    self.step_count, st), "dmic_adam_step")

    def train_step(self, x: torch.Tensor, target: torch.Tensor, lr=1e-3) -> float:
        """One Adam step on mean((net(x) - target/out_scale)^2)."""
        y, acts = self.forward(x, keep_acts=True)
        t = target / self.out_scale
        diff = y - t
        loss = diff.pow(2).mean()
        dy = (2.0 / diff.numel()) * diff
        gw, gb = self.backward(x, acts, dy)
        self.adam(gw, gb, lr=lr)
        return loss


    def direct_step(self, problem: FitProblem, feats: Features, V: torch.Tensor,
                    g: torch.Tensor, adir: torch.Tensor, lr: float = 1e-4,
                    passes: int = 2) -> float:
        """One training step through the meshFT layer: the proposal
        d = smooth(out_scale net(x)) at state (V, grad, Adam direction) is
        applied, and the network descends log L(V + d), the fit loss after
        the jump.  dL/dd comes from the meshFT adjoint (one forward and one
        adjoint of the HIP kernels at the candidate), then through the
        smoothing's transpose and the MLP backward.  Returns L(V + d)."""
        x = feats(V, g, adir)
        y, acts = self.forward(x, keep_acts=True)
        d = feats.smooth(self.out_scale * y, passes)
        Lc, gc = problem.loss_and_grad(V + d)
        Lc_f = float(Lc)
        dy = feats.smooth_adjoint(gc, passes) * (self.out_scale / max(Lc_f, 1e-30))
        gw, gb = self.backward(x, acts, dy)
        self.adam(gw, gb, lr=lr)
        return Lc_f


# ------------------------------------------------------------------ training data
@dataclass
class TrajectoryData:
    x: torch.Tensor        # [rows, 15] features
    target: torch.Tensor   # [rows, 3] displacement over the horizon


def trajectory_states(problem: FitProblem, steps: int, every: int, lr: float = 1e-3):
    """(vertices, gradient, Adam direction, loss) every ``every`` steps of a plain Adam fit."""
    from .fit import fit_adam
    res = fit_adam(problem, steps=steps, lr=lr, snapshot_every=every, sync_every=0)
    return [(V, g, adir, res.losses[k]) for k, V, g, adir in res.snapshots]


def trajectory_samples(problem: FitProblem, feats: Features, steps: int, every: int,
                       horizon: int, lr: float = 1e-3) -> TrajectoryData:
    """Runs a plain Adam fit and turns its snapshots into (features, displacement)
    pairs: the displacement over the next ``horizon`` steps, or to the end of
    the fit when ``horizon`` <= 0."""
    from .fit import fit_adam
    res = fit_adam(problem, steps=steps, lr=lr, snapshot_every=every, sync_every=0)
    by_step = {k: V for k, V, _, _ in res.snapshots}
    by_step[steps] = res.verts
    xs, ts = [], []
    for k, V, g, adir in res.snapshots:
        k2 = steps if horizon <= 0 else min(k + horizon, steps)
        if k2 == k or k2 not in by_step:
            continue
        xs.append(feats(V, g, adir))
        ts.append(by_step[k2] - V)
    return TrajectoryData(torch.cat(xs), torch.cat(ts))


# ------------------------------------------------------------------ learned fit
def _adam_predicted_decrease(opt: torch.optim.Adam, p: torch.Tensor, g: torch.Tensor,
                             adir: torch.Tensor) -> float:
    """First-order loss change g . delta of the Adam step opt would take with gradient g."""
    return float(-(g * adir).sum()) * opt.param_groups[0]["lr"]


@torch.no_grad()
def _adam_absorb(opt: torch.optim.Adam, p: torch.Tensor, g: torch.Tensor) -> None:
    """Advance Adam's moment estimates with gradient g without moving p (the
    iteration's step was an accepted proposal instead)."""
    group = opt.param_groups[0]
    b1, b2 = group["betas"]
    st = opt.state[p]
    if not st:
        st["step"] = torch.tensor(0.0)
        st["exp_avg"] = torch.zeros_like(p, memory_format=torch.preserve_format)
        st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.preserve_format)
    st["step"] += 1
    st["exp_avg"].mul_(b1).add_(g, alpha=1 - b1)
    st["exp_avg_sq"].mul_(b2).addcmul_(g, g, value=1 - b2)


def fit_learned(problem: FitProblem, net: ProposalMLP, feats: Features, steps: int = 1000,
                lr: float = 1e-3, stop_loss: float | None = None, scale0: float = 1.0,
                grow: float = 1.5, shrink: float = 0.5, scale_max: float = 1.0,
                patience: int = 4, cooldown: int = 8, cooldown_max: int = 256,
                snapshot_every: int = 0, accept_margin: float = 8.0,
                precond=None, smooth_passes: int | None = 2,
                line_search: int = 0, absorb: bool = True,
                accept_rule: str = "adam") -> FitResult:
    """Accept/reject fit.  Each iteration the network proposes V + s*d (d
    smoothed over the 1-ring; trust scale s grows after an accepted proposal,
    shrinks after a rejection).  The candidate's loss is computed with its
    autograd graph and accepted only if it is lower than the current loss by
    more than ``accept_margin`` times the first-order decrease of the Adam
    step that would otherwise be taken; its gradient then comes from that
    same graph (one forward + one
    adjoint, the cost of an Adam step) and Adam's moments absorb the gradient
    of the point left behind.  Otherwise the candidate is discarded and the
    plain Adam step is taken.  After ``patience`` rejections in a row no
    proposal is made for ``cooldown`` iterations, a pause that doubles (up to
    ``cooldown_max``) each time it recurs without an acceptance in between;
    the trust scale restarts from ``scale0`` after each pause.
    ``smooth_passes`` 1-ring averaging passes are applied to each proposal
    (None: scaled with the mesh, 2 per 2562 vertices, so the smoothing
    length stays the same on finer meshes).  With ``line_search`` > 0 a
    candidate that passes the bar is tried again at twice its scale (one
    more forward each, at most ``line_search`` times) while that lowers the
    loss.  ``absorb=False`` leaves Adam's moments as they were when a
    proposal is accepted.  ``accept_rule="adam"`` replaces the margin test
    by the exact one: the plain Adam step is taken first (one more forward
    per proposal) and the proposal is accepted only if its loss is below
    both the current loss and the Adam step's loss; otherwise the Adam step
    stands and its graph gives the next gradient.  With
    ``precond`` (an optim.LargeStepsPreconditioner) the smoothed proposal is
    further multiplied by (I + lambda L)^-2, which keeps its smooth
    components and damps the rest.
    ``losses[k]`` is the loss after iteration k.  With ``snapshot_every``,
    (k, vertices, gradient, Adam direction, Adam state) is kept every that many
    iterations (on-policy training data, onpolicy_samples)."""
    Vp = problem.init_verts.clone().requires_grad_(True)
    if smooth_passes is None:
        smooth_passes = max(2, round(2 * Vp.shape[0] / 2562))
    opt = torch.optim.Adam([Vp], lr=lr)
    losses, times = [], []
    accepted = proposals = 0
    accepted_at = []
    s = scale0
    misses = 0
    quiet = 0
    pause = cooldown
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L, g = problem.loss_and_grad(Vp)
    snaps = []
    for k in range(steps):
        if snapshot_every and k % snapshot_every == 0:
            st = opt.state.get(Vp, {})
            snaps.append((k, Vp.detach().clone(), g.detach().clone(),
                          adam_direction(opt, Vp, g),
                          {key: (v.clone() if torch.is_tensor(v) else v) for key, v in st.items()}))
        took = False
        if quiet > 0:
            quiet -= 1
        else:
            adir = adam_direction(opt, Vp, g)
            d = feats.smooth(net.propose(feats(Vp.detach(), g, adir)), passes=smooth_passes)
            if precond is not None:
                d = precond.solve(d)
            proposals += 1
            V0 = Vp.detach().clone()
            La = None
            if accept_rule == "adam":
                # the plain Adam step, taken for real: its loss is part of the
                # bar and its graph gives the gradient if it wins; opt.step()
                # also advances Adam's moments with g (what absorbing does)
                Vp.grad = g
                opt.step()
                La = problem.loss(Vp)
                bar = min(float(L), float(La.detach()))
            else:
                bar = float(L) + accept_margin * _adam_predicted_decrease(opt, Vp, g, adir)
            # the candidate at the trust scale; while it passes the bar and
            # improves, up to line_search more at twice the previous scale
            cand = (V0 + s * d).requires_grad_(True)
            Lc = problem.loss(cand)
            sc = s
            for _ in range(line_search):
                if float(Lc.detach()) >= bar or sc >= scale_max:
                    break
                sc2 = min(2.0 * sc, scale_max)
                c2 = (V0 + sc2 * d).requires_grad_(True)
                l2 = problem.loss(c2)
                if float(l2.detach()) >= float(Lc.detach()):
                    break
                cand, Lc, sc = c2, l2, sc2
            s = sc
            if float(Lc.detach()) < bar:
                (gc,) = torch.autograd.grad(Lc, cand)
                if La is None and absorb:
                    _adam_absorb(opt, Vp, g)
                del La
                with torch.no_grad():
                    Vp.copy_(cand.detach())
                L, g = Lc.detach(), gc
                accepted += 1
                accepted_at.append(k)
                took = True
                misses = 0
                pause = cooldown
                s = min(s * grow, scale_max)
            else:
                del Lc, cand
                s = max(s * shrink, 1.0 / 64)
                misses += 1
                if misses >= patience:
                    quiet, misses = pause, 0
                    pause = min(2 * pause, cooldown_max)
                    s = scale0  # the next proposal starts again from the full step
                if La is not None:  # the Adam step already taken wins
                    (g,) = torch.autograd.grad(La, Vp)
                    L = La.detach()
                    took = True
        if not took:
            Vp.grad = g
            opt.step()
            L, g = problem.loss_and_grad(Vp)
        losses.append(float(L))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if stop_loss is not None and losses[-1] <= stop_loss:
            break
    return FitResult(verts=Vp.detach(), losses=losses, times=times, accepted=accepted,
                     proposals=proposals, accepted_at=accepted_at, snapshots=snaps)


def onpolicy_samples(problem: FitProblem, net: ProposalMLP, feats: Features, steps: int,
                     every: int, horizon: int, lr: float = 1e-3) -> TrajectoryData:
    """Training pairs from the states the learned fit itself visits: the fit
    runs with ``net``; from each snapshot (vertices and Adam state) plain Adam
    continues for ``horizon`` steps and its displacement is the target."""
    res = fit_learned(problem, net, feats, steps=steps, lr=lr, snapshot_every=every)
    xs, ts = [], []
    for k, V, g, adir, st in res.snapshots:
        Vc = V.clone().requires_grad_(True)
        opt = torch.optim.Adam([Vc], lr=lr)
        if st:
            opt.state[Vc] = {key: (v.clone() if torch.is_tensor(v) else v) for key, v in st.items()}
        for _ in range(horizon):
            opt.zero_grad(set_to_none=True)
            problem.loss(Vc).backward()
            opt.step()
        xs.append(feats(V, g, adir))
        ts.append(Vc.detach() - V)
    return TrajectoryData(torch.cat(xs), torch.cat(ts))
