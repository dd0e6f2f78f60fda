"""Learned-proposal fit vs plain Adam (config 5), end to end on one GPU.

    python -m deltamic_amd.learned_fit --train-fits 24 --train-steps 300 \
        --eval-shapes 4 --eval-steps 1000 --out results/learned

1. Trajectories: ``--train-fits`` seeded config-4 fits (target seeds
   4000 + k) run plain Adam for ``--train-steps`` steps; every ``--every``
   steps the features and the displacement over the next ``--horizon`` steps
   are recorded.
2. Training: the proposal MLP (HIP kernels) is trained on those pairs.
3. Evaluation on held-out target seeds (4000 + 10000 + k): plain Adam for
   ``--eval-steps`` steps sets the reference final loss and wall time; the
   accept/reject learned fit runs until it reaches that loss.  The ratio of
   wall times is the reported speed-up.

Writes ``summary.json`` (per shape: losses, times, acceptance) and the
network weights (``proposal.pt``) under ``--out``.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np
import torch

from . import synthetic as syn
from .fit import fit_adam, make_problem
from .proposal import (Features, ProposalMLP, fit_learned, onpolicy_samples, trajectory_samples,
                       trajectory_states)


def log(*a):
    print(*a, flush=True)


def collect(args, device):
    """Features with gradients scaled by nv (grad_scale = 1) and their targets."""
    xs, ts = [], []
    for k in range(args.train_fits):
        t0 = time.perf_counter()
        w = syn.config4(k, subdivisions=args.subdiv, box=args.box)
        prob = make_problem(w, device)
        feats = Features(w.faces, len(w.extra["init_verts"]), prob.lap, device, grad_scale=1.0)
        d = trajectory_samples(prob, feats, args.train_steps, args.every, args.horizon)
        xs.append(d.x)
        ts.append(d.target)
        log(f"trajectory {k + 1}/{args.train_fits}: {d.x.shape[0]} rows "
            f"({time.perf_counter() - t0:.1f} s)")
    x = torch.cat(xs)
    # calibrate the gradient feature scale to unit RMS over the training set
    rms = float(x[:, 6:9].pow(2).sum(1).mean().sqrt())
    gs = 1.0 / max(rms, 1e-30)
    x[:, 6:12] *= gs  # the gradient and its 1-ring mean
    return x, torch.cat(ts), gs


def collect_onpolicy(args, net, device):
    """On-policy rows: the learned fit with ``net`` on the training seeds,
    plain-Adam continuations of ``--horizon`` steps as targets."""
    xs, ts = [], []
    for k in range(args.onpolicy_fits):
        t0 = time.perf_counter()
        w = syn.config4(k, subdivisions=args.subdiv, box=args.box)
        prob = make_problem(w, device)
        feats = Features(w.faces, len(w.extra["init_verts"]), prob.lap, device,
                         grad_scale=net.grad_scale)
        d = onpolicy_samples(prob, net, feats, args.train_steps, args.every, args.horizon)
        xs.append(d.x)
        ts.append(d.target)
        log(f"on-policy fit {k + 1}/{args.onpolicy_fits}: {d.x.shape[0]} rows "
            f"({time.perf_counter() - t0:.1f} s)")
    return torch.cat(xs), torch.cat(ts)


def collect_states(args, device, grad_scale):
    """Training states for the direct objective: per training fit (seed
    4000 + k, mesh subdivision s from --direct-fits "s:count,...") the fit
    problem, its feature builder and the plain-Adam states every --every steps."""
    fits = []
    for spec in args.direct_fits.split(","):
        subdiv, count = (int(v) for v in spec.split(":"))
        for k in range(count):
            t0 = time.perf_counter()
            w = syn.config4(k, subdivisions=subdiv, box=args.box)
            prob = make_problem(w, device)
            feats = Features(w.faces, len(w.extra["init_verts"]), prob.lap, device,
                             grad_scale=grad_scale)
            states = trajectory_states(prob, args.train_steps, args.every)
            fits.append((prob, feats, states, len(w.faces)))
            log(f"states: fit {k + 1}/{count} ({len(w.faces)} faces): {len(states)} states "
                f"({time.perf_counter() - t0:.1f} s)")
    return fits


def train_direct(net, fits, epochs, lr, seed, passes):
    """Descends the mean log L(V + d) over all states (one state per network
    step, shuffled each epoch); logs the mean log-ratio L(V + d) / L(V) per
    face count, the fraction by which one proposal lowers the loss."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    items = [(i, j) for i, (_, _, st, _) in enumerate(fits) for j in range(len(st))]
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(len(items), generator=g).tolist()
        acc = {}
        t0 = time.perf_counter()
        for p in perm:
            i, j = items[p]
            prob, feats, states, nf = fits[i]
            V, gr, adir, L0 = states[j]
            Lc = net.direct_step(prob, feats, V, gr, adir, lr=lr, passes=passes)
            a = acc.setdefault(nf, [0.0, 0])
            a[0] += math.log(max(Lc, 1e-30) / max(L0, 1e-30))
            a[1] += 1
        rec = {str(nf): a[0] / a[1] for nf, a in acc.items()}
        hist.append(rec)
        log(f"direct epoch {ep + 1}/{epochs} ({time.perf_counter() - t0:.1f} s): mean log "
            f"L(V+d)/L(V) " + ", ".join(f"{k} faces {v:+.4f}" for k, v in rec.items()))
    return hist


def train(net, x, t, epochs, batch, lr, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = x.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(x.device)
        tot = 0.0
        nb = 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            tot += float(net.train_step(x[idx].contiguous(), t[idx].contiguous(), lr=lr))
            nb += 1
        hist.append(tot / max(nb, 1))
        log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f}")
    return hist


def evaluate(net, args, device, subdiv, k):
    w = syn.config4(10000 + k, subdivisions=subdiv, box=args.box)
    prob = make_problem(w, device)
    feats = Features(w.faces, len(w.extra["init_verts"]), prob.lap, device,
                     grad_scale=net.grad_scale)
    # warm the band cache and kernels outside the timed fits
    prob.loss_and_grad(prob.init_verts)
    torch.cuda.synchronize()
    adam = fit_adam(prob, steps=args.eval_steps, sync_every=1)
    target = adam.losses[-1]
    # Plain Adam's loss is not always monotone (on some shapes it dips below
    # its step-1000 value and rises again), so its time to the target is the
    # time it first reached it, not its full run.  losses[i] is the loss after
    # i steps, known once step i's forward has run (times[i]).
    first = next(i for i, v in enumerate(adam.losses) if v <= target)
    t_adam = adam.times[first]
    t_adam_full = adam.times[-1]
    precond, t_pre = None, 0.0
    if args.smooth_lambda > 0:
        # built per topology; its time is charged to the learned fit below
        from .optim import LargeStepsPreconditioner
        t0 = time.perf_counter()
        precond = LargeStepsPreconditioner(w.faces, len(w.extra["init_verts"]), args.smooth_lambda,
                                           device)
        torch.cuda.synchronize()
        t_pre = time.perf_counter() - t0
    learned = fit_learned(prob, net, feats, steps=args.eval_steps, stop_loss=target, precond=precond,
                          patience=args.patience, cooldown=args.cooldown,
                          cooldown_max=args.cooldown_max, accept_margin=args.accept_margin,
                          line_search=args.line_search, accept_rule=args.accept_rule,
                          scale_max=args.scale_max,
                          smooth_passes=args.smooth_passes if args.smooth_passes > 0 else None)
    reached = bool(learned.losses and learned.losses[-1] <= target)
    t_learned = (learned.times[-1] + t_pre) if learned.times else float("nan")
    rec = {
        "faces": int(len(w.faces)), "seed": w.seed, "adam_final_loss": target,
        "adam_wall_s": t_adam, "adam_first_reach_step": first,
        "adam_full_wall_s": t_adam_full, "adam_steps": args.eval_steps,
        "learned_reached": reached, "learned_wall_s": t_learned,
        "learned_iters": len(learned.losses), "learned_final_loss": learned.losses[-1],
        "accepted": learned.accepted, "proposals": learned.proposals,
        "speedup": (t_adam / t_learned) if reached else None,
        "adam_loss_curve": adam.losses[::50], "learned_loss_curve": learned.losses[::10],
    }
    log(f"eval faces={rec['faces']} seed={w.seed}: adam {t_adam:.2f} s (step {first}) -> "
        f"{target:.6f}; learned "
        f"{t_learned:.2f} s ({rec['learned_iters']} it, {learned.accepted} accepted) "
        f"speedup {rec['speedup']}")
    return rec


def speedup_summary(evals):
    """Medians of the speed-up: over the shapes that reached plain Adam's loss
    (median_speedup), and over every shape with a miss counted as its
    (pessimistic) ratio t_adam / t_learned_full (median_speedup_all), overall
    and per face count."""
    def stats(es):
        sp = [e["speedup"] for e in es if e["speedup"] is not None]
        allr = [e["speedup"] if e["speedup"] is not None else e["adam_wall_s"] / e["learned_wall_s"]
                for e in es]
        return {"median_speedup": float(np.median(sp)) if sp else None,
                "median_speedup_all": float(np.median(allr)) if allr else None,
                "reached": f"{len(sp)}/{len(es)}",
                "at_least_3x": int(sum(1 for r in allr if r >= 3.0))}
    out = stats(evals)
    out["by_faces"] = {str(f): stats([e for e in evals if e["faces"] == f])
                       for f in sorted({e["faces"] for e in evals})}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-fits", type=int, default=24)
    ap.add_argument("--train-steps", type=int, default=300)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--horizon", type=int, default=100,
                    help="displacement target over this many steps (<= 0: to the end of the fit)")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-shapes", type=int, default=4)
    ap.add_argument("--eval-shapes-large", type=int, default=1)
    ap.add_argument("--eval-steps", type=int, default=1000)
    ap.add_argument("--eval-offset", type=int, default=0,
                    help="first held-out shape index (target seeds 14000 + offset + k, and "
                         "14100 + offset + k for the larger meshes)")
    ap.add_argument("--subdiv", type=int, default=4)
    ap.add_argument("--box", type=int, default=128)
    ap.add_argument("--out", default="gpurun_out/learned")
    ap.add_argument("--load", default=None, help="skip training; load proposal.pt")
    ap.add_argument("--no-eval", action="store_true", help="train and save only")
    ap.add_argument("--onpolicy-fits", type=int, default=0,
                    help="after training (or --load), run the learned fit on this many training "
                         "seeds, take plain-Adam continuations from its states as targets, and "
                         "train --onpolicy-epochs more epochs on those rows")
    ap.add_argument("--onpolicy-epochs", type=int, default=20)
    ap.add_argument("--patience", type=int, default=4,
                    help="rejections in a row before a proposal-free pause")
    ap.add_argument("--cooldown", type=int, default=8, help="first pause length (iterations)")
    ap.add_argument("--smooth-lambda", type=float, default=0.0,
                    help="also smooth each proposal by (I + lambda L)^-2 (0: off)")
    ap.add_argument("--accept-margin", type=float, default=8.0,
                    help="accept a proposal only if it beats this many times the first-order "
                         "decrease of the Adam step it replaces")
    ap.add_argument("--cooldown-max", type=int, default=256,
                    help="pauses double up to this length")
    ap.add_argument("--direct-fits", default="",
                    help="train through the meshFT layer (descend log L(V + d)) on plain-Adam "
                         "states of these fits, 'subdivision:count,...' (e.g. '4:24,5:8'), "
                         "starting from --load's weights or a fresh network")
    ap.add_argument("--direct-epochs", type=int, default=10)
    ap.add_argument("--direct-lr", type=float, default=1e-4)
    ap.add_argument("--accept-rule", choices=("adam", "margin"), default="adam",
                    help="adam: accept a proposal only if it beats the plain Adam step's actual "
                         "loss (and the current loss); margin: beat --accept-margin times the "
                         "Adam step's first-order decrease")
    ap.add_argument("--scale-max", type=float, default=1.0,
                    help="largest trust scale applied to the network's proposal")
    ap.add_argument("--line-search", type=int, default=0,
                    help="fit_learned: doublings of an accepted proposal's scale while they help")
    ap.add_argument("--smooth-passes", type=int, default=2,
                    help="1-ring smoothing passes of each proposal (0: scaled with the mesh)")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    device = torch.device("cuda:0")
    net = ProposalMLP(device, seed=0)
    summary = {"args": vars(args)}
    if args.load:
        net.load_state_dict(torch.load(args.load, weights_only=True))
    if args.direct_fits:
        if net.grad_scale is None:
            raise SystemExit("--direct-fits needs --load weights (their gradient feature scale)")
        t0 = time.perf_counter()
        fits = collect_states(args, device, net.grad_scale)
        summary["direct_states"] = {str(nf): sum(len(f[2]) for f in fits if f[3] == nf)
                                    for nf in sorted({f[3] for f in fits})}
        summary["direct_collect_s"] = time.perf_counter() - t0
        net.step_count = 0  # fresh Adam moments for the network's own optimiser
        net.mw.zero_(); net.vw.zero_(); net.mb.zero_(); net.vb.zero_()
        t1 = time.perf_counter()
        passes = args.smooth_passes if args.smooth_passes > 0 else 2
        summary["direct_train"] = train_direct(net, fits, args.direct_epochs, args.direct_lr,
                                               seed=3, passes=passes)
        summary["direct_train_s"] = time.perf_counter() - t1
        torch.save(net.state_dict(), os.path.join(args.out, "proposal.pt"))
        del fits
    elif not args.load:
        t0 = time.perf_counter()
        x, t, gs = collect(args, device)
        net.grad_scale = gs
        summary["grad_scale"] = gs
        summary["train_rows"] = int(x.shape[0])
        summary["collect_s"] = time.perf_counter() - t0
        t1 = time.perf_counter()
        summary["train_loss"] = train(net, x, t, args.epochs, args.batch, args.lr, seed=1)
        summary["train_s"] = time.perf_counter() - t1
        torch.save(net.state_dict(), os.path.join(args.out, "proposal.pt"))
    if args.onpolicy_fits > 0:
        t0 = time.perf_counter()
        xo, to = collect_onpolicy(args, net, device)
        summary["onpolicy_rows"] = int(xo.shape[0])
        summary["onpolicy_collect_s"] = time.perf_counter() - t0
        net.step_count = 0  # fresh Adam moments for the network's own optimiser
        net.mw.zero_(); net.vw.zero_(); net.mb.zero_(); net.vb.zero_()
        t1 = time.perf_counter()
        summary["onpolicy_train_loss"] = train(net, xo, to, args.onpolicy_epochs, args.batch,
                                               args.lr, seed=2)
        summary["onpolicy_train_s"] = time.perf_counter() - t1
        torch.save(net.state_dict(), os.path.join(args.out, "proposal.pt"))
    with open(os.path.join(args.out, "summary.json"), "w") as f:
        json.dump(summary, f, indent=1)
    if args.no_eval:
        return
    evals = []
    o = args.eval_offset
    jobs = [(args.subdiv, o + k) for k in range(args.eval_shapes)]
    jobs += [(args.subdiv + 1, 100 + o + k) for k in range(args.eval_shapes_large)]
    for subdiv, k in jobs:
        evals.append(evaluate(net, args, device, subdiv, k))
        summary["eval"] = evals
        summary.update(speedup_summary(evals))
        with open(os.path.join(args.out, "summary.json"), "w") as f:  # after every shape
            json.dump(summary, f, indent=1)
    log(json.dumps({k: summary[k] for k in ("median_speedup", "median_speedup_all", "reached",
                                            "by_faces")}))


if __name__ == "__main__":
    main()
