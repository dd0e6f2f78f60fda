#!/usr/bin/env python3
"""meshFT forward + adjoint throughput on one MI355X.

Workload (one "step"): the 20,480-face icosphere (subdivision 5, radial
spherical-harmonic displacement l <= 6, c ~ N(0, 0.03), R = 1, seed 2000) in a
200^3 box under a Gaussian OTF Sigma = 1e-3 I with band threshold 1e-5
(1,081,808 band frequencies): one meshFT forward and one adjoint through the C-ABI
on inputs already resident in HBM (vertices, faces, band frequencies, an
upstream gradient).  value = faces * band / step time, in G face.freq/s.

The metric's second half, the learned-proposal fit's wall time to plain
Adam's loss against plain Adam's, is the `learned_fit` leg (one held-out
shape with 5120 and one with 20480 faces, 128^3).

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "meshFT fwd+bwd Gface·freq/s; fit time to loss tol, learned vs plain Adam"
UNIT = "Gface·freq/s"
FP32_PEAK_TFLOPS = 157.3
FLOPS_FWD = 20   # algorithmic FP32 flops per face.frequency pair, forward
FLOPS_BWD = 85   # adjoint


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """(threads to use, description of the host): the threads are this GPU's
    share of the host (OMP_NUM_THREADS, else the CPUs this process may run on)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = int(omp) if omp and omp.isdigit() and int(omp) > 0 else affinity
    threads = max(1, min(threads, affinity))
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": omp,
                     "model": model}


def cpu_baseline(w, xi_np, Vs, n_sample: int):
    """Reference-semantics FP32 restatement (oracle), fwd + adjoint, on this
    GPU's share of the host cores."""
    from oracle import oracle as O
    from deltamic_amd import synthetic as syn
    cores, host = host_cpu()
    sel = np.sort(syn.rng(w.seed + 1).choice(len(xi_np), n_sample, replace=False))
    xi = xi_np[sel]
    g = syn.rng(w.seed + 4)
    grad = (g.standard_normal(n_sample) + 1j * g.standard_normal(n_sample)).astype(np.complex64)
    t0 = time.perf_counter()
    ft = O.ref_forward(Vs, w.faces, w.coeff, xi, np.float32, threads=cores)
    t1 = time.perf_counter()
    O.ref_backward(Vs, w.faces, w.coeff, xi, grad, ft, np.float32, threads=cores)
    t2 = time.perf_counter()
    pairs = len(w.faces) * n_sample
    return {
        "value": pairs / (t2 - t0) / 1e9,
        "unit": UNIT,
        "cores": cores,
        "host": host,
        "kind": "port",
        "sample": (f"same mesh, {n_sample} of the band frequencies (seed {w.seed + 1}); "
                   f"reference-semantics FP32 forward {t1 - t0:.2f} s + adjoint {t2 - t1:.2f} s "
                   f"(oracle/meshft_oracle.cpp, g++ -O3 -fopenmp)"),
    }


def psf_leg(L, dev, w, steps):
    """Learnable Gaussian PSF on the bench box: forward (writes 4 B/voxel) and
    sigma gradient (reads 8 B/voxel), timed with events on the launch stream."""
    import torch
    from deltamic_amd import _lib
    n0, n1, n2 = (int(x) for x in w.box_shape)
    axes = [torch.linspace(-1.0, 1.0, n, device=dev) for n in (n0, n1, n2)]
    S = (1e3 * torch.eye(3, device=dev)).contiguous()
    out = torch.empty((n0, n1, n2), device=dev)
    G = torch.ones_like(out)
    gS = torch.empty(9, device=dev)
    nbytes = int(L.dmic_psf_grad_workspace(n0, n1, n2))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    st = _lib.stream_of(out)
    ap = [a.data_ptr() for a in axes]

    def fwd():
        _lib.check(L.dmic_psf_gaussian(S.data_ptr(), *ap, n0, n1, n2, out.data_ptr(), st),
                   "dmic_psf_gaussian")

    def grad():
        _lib.check(L.dmic_psf_gaussian_grad(G.data_ptr(), out.data_ptr(), *ap, n0, n1, n2,
                                            gS.data_ptr(), ws.data_ptr(), nbytes, st),
                   "dmic_psf_gaussian_grad")

    res = {}
    for name, fn, nbytes_moved in (("forward", fwd, 4 * n0 * n1 * n2),
                                   ("grad", grad, 8 * n0 * n1 * n2)):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        res[f"{name}_ms"] = round(ms, 4)
        res[f"{name}_GBps"] = round(nbytes_moved / (ms * 1e-3) / 1e9, 1)
    res["bytes_forward"] = 4 * n0 * n1 * n2
    res["bytes_grad"] = 8 * n0 * n1 * n2
    res["hbm_peak_GBps"] = 8000.0
    return res


def render_leg(dev, w, band, otf, vals, reps):
    """The render chain around meshFT on the bench box, op by op (render.py:5-6,
    177-180; cuda_class.py:27-28,36): each op's average time from events on
    the stream it runs on, its algorithmic HBM bytes (each operand read once,
    each result written once) and the rate against 8 TB/s."""
    import torch
    from deltamic_amd import normalize_tensor, render_image_from_ftmesh
    from deltamic_amd.hip_class import gather_band, scatter_band
    nbox = int(np.prod(w.box_shape))
    n = band.idx.numel()
    shape = list(w.box_shape)
    box = scatter_band(vals, band)
    prod = box * otf
    spec = torch.fft.ifftn(prod, shape)
    img = torch.abs(spec)
    up = torch.randn(nbox, device=dev).reshape(shape)
    from deltamic_amd.render import _AbsHIP
    img_h = _AbsHIP.apply(spec)
    ops = [
        # name, fn, algorithmic bytes (each operand read once, each result written once)
        ("scatter_band", lambda: scatter_band(vals, band), 8 * nbox + 12 * n),
        ("otf_multiply", lambda: box * otf, 20 * nbox),
        ("ifftn", lambda: torch.fft.ifftn(prod, shape), 16 * nbox),
        ("abs_minmax", lambda: _AbsHIP.apply(spec), 12 * nbox),
        ("normalize", lambda: normalize_tensor(img_h), 16 * nbox),
        ("torch_abs", lambda: torch.abs(spec), 12 * nbox),
        ("torch_normalize", lambda: ((img - img.min()) / (img.max() - img.min()) + 0) / 1,
         16 * nbox),
        ("gather_band", lambda: gather_band(prod, band), 12 * n),
    ]
    res = {}
    for name, fn, nbytes in ops:
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[name] = {"ms": round(ms, 4), "bytes": nbytes,
                     "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    # the whole chain's backward (normalize, abs, ifftn, multiply) to the ftmesh box
    boxg = box.clone().requires_grad_(True)

    def chain_bwd():
        out = normalize_tensor(render_image_from_ftmesh(boxg, otf, shape))
        (gb,) = torch.autograd.grad(out, boxg, up)
        return gb

    def chain_bwd_torch():
        t = torch.abs(torch.fft.ifftn(boxg * otf, shape))
        out = (t - t.min()) / (t.max() - t.min())
        (gb,) = torch.autograd.grad(out, boxg, up)
        return gb

    for key, fn in (("chain_forward_and_backward_ms", chain_bwd),
                    ("torch_chain_forward_and_backward_ms", chain_bwd_torch)):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        res[key] = round(a.elapsed_time(b) / reps, 4)
    res["box"] = list(w.box_shape)
    res["band"] = n
    res["hbm_peak_GBps"] = 8000.0
    return res


def closure_leg(dev, w, Vs, reps):
    """The reference fit loop's OTF (Gradient_Descent.ipynb cell 2): OTF =
    |fftshift(fftn(PSF))| of Gaussian_psf(box_shape, box_size[:, 1], sigma=1e3),
    band threshold 1e-2, on the bench mesh and box.  That band is not
    centrosymmetric (fftshift centres it on k = n/2, the grid mirror is
    k <-> n-1-k), so it is evaluated through its symmetric closure; timed
    against evaluating every band frequency (forward + adjoint each)."""
    import torch
    from deltamic_amd import hip_class
    from deltamic_amd.fourier3d_cpp import backward_fourier3d_cpp, forward_fourier3d_cpp
    from deltamic_amd.psf import Gaussian_psf
    from deltamic_amd.render import spatial_frequency_axes
    bs = torch.tensor(w.box_size, dtype=torch.float32, device=dev)
    axes = spatial_frequency_axes(bs, w.box_shape, torch.float32, dev)
    psf = Gaussian_psf(w.box_shape, torch.tensor(w.box_size)[:, 1], sigma=1e3, device=dev)
    with torch.no_grad():
        otf = torch.abs(torch.fft.fftshift(torch.fft.fftn(psf.forward()))).contiguous()
    band = hip_class.compute_band(otf, 1e-2, *axes)
    V = Vs.to(dev)
    Fc = torch.tensor(w.faces, device=dev)
    C = torch.tensor(w.coeff, device=dev)
    g = torch.randn(band.n, dtype=torch.complex64, device=dev)

    def full():
        ft = forward_fourier3d_cpp(V, Fc, C, band.xi_list)
        backward_fourier3d_cpp(V, Fc, C, g, ft, band.xi_list)

    def closed():
        ft = hip_class.band_forward(V, Fc, C, band)
        hip_class.band_backward(V, Fc, C, g, ft, band)

    res = {"band": band.n, "symmetric": band.symmetric,
           "closure": band.closure.n if band.closure is not None else None}
    for name, fn in (("full_band", full), ("closure_halving", closed)):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[name] = {"ms_per_step": round(ms, 4),
                     "value": round(len(w.faces) * band.n / (ms * 1e-3) / 1e9, 3)}
    return res


def learned_fit_leg(dev, steps):
    """The second half of the metric: wall time of the learned-proposal fit
    to plain Adam's loss, against plain Adam's (lr 1e-3, MSE + 0.1 uniform
    Laplacian, 128^3, Gaussian OTF 1e-3 I, Poisson + Gaussian noise), on one
    held-out target shape of each size (seeds 14500 at 5120 faces and 14600
    at 20480 faces; no training fit used them).  Plain Adam runs `steps`
    steps; its time is the wall time at which it first reached its final
    loss; the learned fit (shipped weights, default policy) runs until it
    reaches that loss."""
    import torch
    from deltamic_amd import synthetic as syn
    from deltamic_amd.fit import fit_adam, make_problem
    from deltamic_amd.proposal import N_IN, Features, ProposalMLP, fit_learned
    net = ProposalMLP(dev)
    net.load_state_dict(torch.load(os.path.join(ROOT, "deltamic_amd", "weights", "proposal.pt"),
                                   weights_only=True))
    res = {"adam_steps": steps, "shapes": []}
    for k, subdiv in ((10500, 4), (10600, 5)):
        w = syn.config4(k, subdivisions=subdiv)
        prob = make_problem(w, dev)
        feats = Features(w.faces, prob.init_verts.shape[0], prob.lap, dev,
                         grad_scale=net.grad_scale)
        prob.loss_and_grad(prob.init_verts)
        torch.cuda.synchronize()
        adam = fit_adam(prob, steps=steps)
        target = adam.losses[-1]
        first = next(i for i, v in enumerate(adam.losses) if v <= target)
        learned = fit_learned(prob, net, feats, steps=steps, stop_loss=target)
        reached = learned.losses[-1] <= target
        t_l = learned.times[-1]
        rec = {"faces": len(w.faces), "seed": w.seed, "adam_final_loss": round(target, 7),
               "adam_first_reach_step": first, "adam_s": round(adam.times[first], 3),
               "learned_iterations": len(learned.losses), "learned_accepted": learned.accepted,
               "learned_s": round(t_l, 3), "reached": bool(reached),
               "speedup": round(adam.times[first] / t_l, 3)}
        # the proposal network alone: forward (keeping activations) + backward
        # over one row per vertex, 20 repetitions between HIP events
        nv = prob.init_verts.shape[0]
        x = torch.randn(nv, N_IN, device=dev, generator=torch.Generator(dev).manual_seed(5))
        y, acts = net.forward(x, keep_acts=True)
        net.backward(x, acts, y)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(20):
            y, acts = net.forward(x, keep_acts=True)
            net.backward(x, acts, y)
        b.record()
        torch.cuda.synchronize()
        rec["mlp_forward_backward_ms"] = round(a.elapsed_time(b) / 20, 4)
        rec["mlp_rows"] = nv
        res["shapes"].append(rec)
        log(f"learned fit, {len(w.faces)} faces (seed {w.seed}): plain Adam first reaches its "
            f"{steps}-step loss {target:.6f} at step {first} ({adam.times[first]:.2f} s); learned "
            f"fit {len(learned.losses)} iterations {t_l:.2f} s ({'reached' if reached else 'MISSED'}"
            f"): {rec['speedup']:.2f}x; proposal MLP forward+backward over {nv} rows "
            f"{rec['mlp_forward_backward_ms']} ms")
        del prob, feats
    res["min_speedup"] = min(r["speedup"] for r in res["shapes"])
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-freqs", type=int, default=2048,
                    help="band frequencies in the CPU-baseline sample (0 = skip)")
    ap.add_argument("--no-halving", action="store_true",
                    help="evaluate every band frequency instead of mirroring half of the band")
    ap.add_argument("--full-band-steps", type=int, default=3,
                    help="also time this many steps evaluating every band frequency")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB three times (reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--module-steps", type=int, default=3,
                    help="extra end-to-end module steps (forward, render, backward) to time")
    ap.add_argument("--af-steps", type=int, default=3,
                    help="also time the autofluorescence variant (forward + adjoint) on the "
                         "same mesh and band (reported under 'autofluorescence')")
    ap.add_argument("--render-reps", type=int, default=10,
                    help="time each op of the render chain this many times (0 = skip; "
                         "reported under 'render')")
    ap.add_argument("--closure-steps", type=int, default=3,
                    help="time the fit loop's |fftshift(fftn(PSF))| OTF band through its "
                         "symmetric closure and in full (0 = skip; reported under 'fftshift_otf')")
    ap.add_argument("--fit-steps", type=int, default=1000,
                    help="plain-Adam steps of the learned-fit leg (0 = skip; reported under "
                         "'learned_fit')")
    ap.add_argument("--psf-steps", type=int, default=5,
                    help="also time the learnable-PSF kernels (forward and sigma gradient) on "
                         "the box (HBM-bound; reported separately under 'psf')")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs meshFT on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import torch
    from deltamic_amd import (Fourier3dMesh, generate_gaussian_psf, normalize_tensor,
                              render_image_from_ftmesh, _lib)
    from deltamic_amd import synthetic as syn
    from deltamic_amd.hip_class import compute_band

    dev = torch.device("cuda:0")
    w = syn.config2()
    t0 = time.perf_counter()
    box = torch.tensor(w.box_size)
    probe = Fourier3dMesh(box, w.box_shape, device=dev, dtype=torch.float32)
    otf = generate_gaussian_psf(torch.tensor(w.sigma, device=dev), probe.xi0, probe.xi1, probe.xi2)
    band = compute_band(otf, w.thresh, probe.xi0, probe.xi1, probe.xi2)
    n = band.idx.numel()
    nf, nv = len(w.faces), len(w.verts)
    Vs = (torch.tensor(w.verts) - probe.box_size.cpu()[:, 0]).contiguous()
    V = Vs.to(dev)
    Fc = torch.tensor(w.faces, device=dev)
    C = torch.tensor(w.coeff, device=dev)
    g = syn.rng(w.seed + 4)
    grad = torch.tensor((g.standard_normal(n) + 1j * g.standard_normal(n)).astype(np.complex64),
                        device=dev)
    out = torch.empty((n,), dtype=torch.complex64, device=dev)
    gv = torch.empty((nv, 3), dtype=torch.float32, device=dev)
    L = _lib.lib()
    # Hermitian halving (the band of this grid and OTF is exactly centrosymmetric):
    # the kernels evaluate the first (n+1)/2 frequencies, the rest is mirrored.
    halve = bool(band.symmetric) and not args.no_halving
    ne = band.half if halve else n
    g_half = torch.empty((ne,), dtype=torch.complex64, device=dev)
    wsf = _lib.workspace(L.dmic_meshft_forward_workspace(nv, nf, ne), dev)
    wsb = _lib.workspace(L.dmic_meshft_backward_workspace(nv, nf, ne), dev)
    wsf_full = _lib.workspace(L.dmic_meshft_forward_workspace(nv, nf, n), dev) if halve else wsf
    wsb_full = _lib.workspace(L.dmic_meshft_backward_workspace(nv, nf, n), dev) if halve else wsb
    st = torch.cuda.current_stream(dev).cuda_stream
    log(f"setup {time.perf_counter() - t0:.1f} s: nv={nv} nf={nf} band={n} "
        f"symmetric={band.symmetric} evaluated={ne}")

    def step_full():
        # every band frequency evaluated (no mirroring): reported beside the headline
        rc = L.dmic_meshft_forward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(),
                                   band.xi_list.data_ptr(), n, out.data_ptr(), wsf_full.data_ptr(),
                                   wsf_full.numel(), st)
        _lib.check(rc, "forward")
        rc = L.dmic_meshft_backward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(),
                                    grad.data_ptr(), out.data_ptr(), band.xi_list.data_ptr(), n,
                                    gv.data_ptr(), wsb_full.data_ptr(), wsb_full.numel(), st)
        _lib.check(rc, "backward")

    def step():
        rc = L.dmic_meshft_forward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(),
                                   band.xi_list.data_ptr(), ne, out.data_ptr(), wsf.data_ptr(),
                                   wsf.numel(), st)
        _lib.check(rc, "forward")
        if halve:
            _lib.check(L.dmic_band_mirror_fill(out.data_ptr(), n, st), "mirror_fill")
            _lib.check(L.dmic_band_mirror_fold(grad.data_ptr(), n, g_half.data_ptr(), st),
                       "mirror_fold")
        gsrc = g_half if halve else grad
        rc = L.dmic_meshft_backward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(),
                                    gsrc.data_ptr(), out.data_ptr(), band.xi_list.data_ptr(), ne,
                                    gv.data_ptr(), wsb.data_ptr(), wsb.numel(), st)
        _lib.check(rc, "backward")

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup}")
    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            # an elementwise kernel (not the DMA engine a plain copy_ may use):
            # reads nbytes, writes nbytes
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log(f"copy calibration: 3 x elementwise kernel reading {nbytes} and writing {nbytes} bytes")
    L.dmic_profile_enable(1)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    t_wall = time.perf_counter()
    for i in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms_total = e0.elapsed_time(e1)
    fwd_ms, fwd_n, bwd_ms, bwd_n = (ctypes.c_double(), ctypes.c_long(), ctypes.c_double(),
                                    ctypes.c_long())
    L.dmic_profile_read(ctypes.byref(fwd_ms), ctypes.byref(fwd_n), ctypes.byref(bwd_ms),
                        ctypes.byref(bwd_n))
    L.dmic_profile_enable(0)
    ms_step = ms_total / args.steps
    pairs = nf * n              # delivered: every band frequency of the full band
    kpairs = nf * ne            # evaluated by each of the two kernels
    value = pairs / (ms_step * 1e-3) / 1e9
    kf = fwd_ms.value / max(fwd_n.value, 1)
    kb = bwd_ms.value / max(bwd_n.value, 1)
    log(f"step {ms_step:.3f} ms (wall {t_wall / args.steps * 1e3:.3f} ms): fwd kernel {kf:.3f} ms, "
        f"bwd kernel {kb:.3f} ms")
    ach_f = FLOPS_FWD * kpairs / (kf * 1e-3) / 1e12
    ach_b = FLOPS_BWD * kpairs / (kb * 1e-3) / 1e12
    dom = ("meshft_bwd_kernel", ach_b, kb) if kb >= kf else ("meshft_fwd_kernel", ach_f, kf)

    full_band = None
    if halve and args.full_band_steps > 0:
        step_full()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.full_band_steps):
            step_full()
        b.record()
        torch.cuda.synchronize()
        fms = a.elapsed_time(b) / args.full_band_steps
        full_band = {"value": round(pairs / (fms * 1e-3) / 1e9, 3), "ms_per_step": round(fms, 4),
                     "steps": args.full_band_steps}
        log(f"full-band evaluation (no mirroring): {fms:.3f} ms/step, {full_band['value']} {UNIT}")

    # end-to-end module step: forward, render (hipFFT), normalize, loss, backward
    module_ms = None
    if args.module_steps > 0:
        mesh_ft = Fourier3dMesh(box, w.box_shape, dev, torch.float32, otf, w.thresh)
        Vp = torch.tensor(w.verts, device=dev, requires_grad=True)

        def mstep():
            Vp.grad = None
            img = normalize_tensor(render_image_from_ftmesh(mesh_ft(Vp, Fc, C), otf, w.box_shape))
            img.sum().backward()

        mstep()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.module_steps):
            mstep()
        b.record()
        torch.cuda.synchronize()
        module_ms = a.elapsed_time(b) / args.module_steps
        log(f"module step (fwd+render+normalize+bwd): {module_ms:.3f} ms")

    af = None
    if args.af_steps > 0:
        wsb_af = _lib.workspace(L.dmic_meshft_af_backward_workspace(nv, nf, ne), dev)

        def step_af():
            rc = L.dmic_meshft_af_forward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(), 0.5,
                                          band.xi_list.data_ptr(), ne, out.data_ptr(),
                                          wsf.data_ptr(), wsf.numel(), st)
            _lib.check(rc, "af forward")
            if halve:
                _lib.check(L.dmic_band_mirror_fill(out.data_ptr(), n, st), "mirror_fill")
                _lib.check(L.dmic_band_mirror_fold(grad.data_ptr(), n, g_half.data_ptr(), st),
                           "mirror_fold")
            gsrc = g_half if halve else grad
            rc = L.dmic_meshft_af_backward(V.data_ptr(), nv, Fc.data_ptr(), 1, nf, C.data_ptr(), 0.5,
                                           gsrc.data_ptr(), band.xi_list.data_ptr(), ne,
                                           gv.data_ptr(), wsb_af.data_ptr(), wsb_af.numel(), st)
            _lib.check(rc, "af backward")

        step_af()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.af_steps):
            step_af()
        b.record()
        torch.cuda.synchronize()
        ams = a.elapsed_time(b) / args.af_steps
        af = {"value": round(pairs / (ams * 1e-3) / 1e9, 3), "ms_per_step": round(ams, 4),
              "steps": args.af_steps, "regint": 0.5}
        log(f"autofluorescence variant: {ams:.3f} ms/step, {af['value']} {UNIT}")

    render = None
    if args.render_reps > 0:
        render = render_leg(dev, w, band, otf, out, args.render_reps)
        log("render chain on " + str(w.box_shape) + ": " + ", ".join(
            f"{k} {v['ms']:.4f} ms ({v['GBps']:.0f} GB/s)" for k, v in render.items()
            if isinstance(v, dict)) + f"; chain fwd+bwd {render['chain_forward_and_backward_ms']} ms "
            f"(torch ops: {render['torch_chain_forward_and_backward_ms']} ms)")

    psf = None
    if args.psf_steps > 0:
        psf = psf_leg(L, dev, w, args.psf_steps)
        log(f"psf kernels on {w.box_shape}: forward {psf['forward_ms']:.4f} ms "
            f"({psf['forward_GBps']:.0f} GB/s), sigma gradient {psf['grad_ms']:.4f} ms "
            f"({psf['grad_GBps']:.0f} GB/s)")

    cpu = None
    if args.cpu_freqs > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(w, band.xi_list.cpu().numpy(), Vs.numpy(), args.cpu_freqs)
        log(f"cpu baseline {cpu['value']:.4f} {UNIT} on {cpu['cores']} threads")

    closure = None
    if args.closure_steps > 0:
        closure = closure_leg(dev, w, Vs, args.closure_steps)
        log(f"fftshift OTF band {closure['band']} (closure {closure['closure']}): full band "
            f"{closure['full_band']['value']} {UNIT}, through the closure "
            f"{closure['closure_halving']['value']} {UNIT}")

    lfit = None
    if args.fit_steps > 0:
        lfit = learned_fit_leg(dev, args.fit_steps)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "meshft_cfg2_pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("band_evaluated_per_kernel") == ne and dom[0] in tr.get("bytes_per_launch", {}):
            traffic = tr["bytes_per_launch"][dom[0]]

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("icosphere s5 (10242 v, 20480 f) + real SH l<=6 c~N(0,0.03), R=1, seed 2000; "
                         "200^3 box (compute_box_size offset 0.2); Gaussian OTF Sigma=1e-3 I; "
                         "narrowband_thresh 1e-5; one meshFT forward + adjoint via the C-ABI"),
            "faces": nf,
            "verts": nv,
            "band": n,
            "box": list(w.box_shape),
            "hermitian_halving": halve,
            "band_evaluated_per_kernel": ne,
        },
        "roofline": {
            "bound": "valu",
            "kernel": dom[0],
            "achieved": round(dom[1], 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(dom[1] / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "flops_per_pair": FLOPS_BWD if dom[0] == "meshft_bwd_kernel" else FLOPS_FWD,
            "pairs_per_launch": kpairs,
            "kernel_ms": round(dom[2], 4),
        },
        "kernels": {
            "meshft_fwd_kernel_ms": round(kf, 4),
            "meshft_bwd_kernel_ms": round(kb, 4),
            "fwd_tflops": round(ach_f, 3),
            "bwd_tflops": round(ach_b, 3),
            "module_step_ms": None if module_ms is None else round(module_ms, 3),
        },
        "full_band_evaluation": full_band,
        "autofluorescence": af,
        "fftshift_otf": closure,
        "learned_fit": lfit,
        "render": render,
        "psf": psf,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
