#!/usr/bin/env python3
"""Benchmark of the coupled 4-channel Frobenius TGV-KL primal-dual
reconstruction on one MI355X.

Workload (one "step" = one primal-dual iteration over the whole volume):
256 x 256 x 64 slices x 4 channels (HAADF-like + 3 EDX-like) seeded
granule phantom, 60 tilts in [-75, 75] degrees, 256 detector bins, Poisson
data, alpha = (4, 1), mu = (0.2, 5e-4, 4e-3, 5e-4), FP32 state, inputs
resident in HBM.  Prints one JSON line with voxel-iters/s, the roofline of
the dominant kernel (HIP events over its launches), the same byte rates for
the two projectors (vector-ALU and LDS bound, far below the HBM peak) and a
CPU baseline (the FP32 C/OpenMP restatement of the reference kernels on this
host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# FP32 vector FMA peak without packed instructions (the library compiles with
# packed FP32 off): 256 CUs x 64 FMA/clk x 2 flops x 2.4 GHz
VALU_PEAK_TFLOPS = 78.6
ALPHA = (4.0, 1.0)
MU = (0.2, 5e-4, 4e-3, 5e-4)
KERNELS = ("dual_single_kernel", "fwd_single_kernel", "adj_single_kernel",
           "primal_single_kernel", "pd_fused_kernel")


def kernel_bytes(prob):
    """Algorithmic HBM bytes per launch of each kernel: every array the kernel
    must touch, read once and written once (neighbour reads not counted)."""
    V = prob.voxels()                       # voxel-channels
    B = prob.C * prob.Nz * prob.na * prob.nd  # bin-channels
    d = 2 if prob.dim == 2 else 3
    m = 3 if prob.dim == 2 else 6
    tgv = prob.regularisation == "TGV"
    # dual step: read ubar, pbar, v, w; write v, w
    dual = V * 4 * (1 + (d if tgv else 0) + 2 * d + (2 * m if tgv else 0))
    # forward projection + prox: read ubar; per bin read lam, f, write lam
    fwd = V * 4 + B * 12
    # backprojection: read lam, write K^T lam
    adj = B * 4 + V * 4
    # primal step: read u, K^T lam, v, w, p; write u, ubar, p, pbar
    primal = V * 4 * (2 + d + 2 + ((m + 3 * d) if tgv else 0))
    # primal step + next dual step in one kernel: read u, K^T lam, v, w, p;
    # write u, ubar, p, v, w (pbar is formed in registers, not stored)
    fused = V * 4 * (2 + d + (m + d if tgv else 0) + 2 + d + (m + d if tgv else 0))
    return dict(zip(KERNELS, (dual, fwd, adj, primal, fused)))


def launch_counts(n, fused):
    """Launches of (dual, fwd, adj, primal, fused) kernels in n iterations of
    gr_tgv_iterate: with the fused form, one unfused iteration when n - 1 is
    odd, then a standalone dual, n - 1 - that fused steps and a standalone
    primal."""
    if not fused or n < 2:
        return (n, n, n, n, 0)
    plain = (n - 1) % 2
    return (plain + 1, n, n, plain + 1, n - plain - 1)


def _native_cref():
    """The FP32 C restatement built for this host's CPU (falls back to the
    prebuilt x86-64-v3 library)."""
    from oracle import cref
    try:
        tmp = tempfile.mkdtemp(prefix="cref_")
        lib_path = os.path.join(tmp, "libcpuref_native.so")
        subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared",
                        "-std=c11", os.path.join(ROOT, "oracle", "cpu_ref.c"), "-o", lib_path,
                        "-lm"], check=True, capture_output=True, timeout=120)
        return lib_path, "native"
    except Exception:
        cref.load()
        return cref._LIB, "x86-64-v3"


def _time_cref(lib_path, f_host, angles, L, iters):
    """Seconds per iteration of the C restatement: the difference of a
    (1 + iters)- and a 1-iteration run divided by iters (allocation and
    set-up excluded)."""
    from oracle import cref
    from oracle import tgv_oracle as O
    lib = cref.load(lib_path)
    C, Nz, na, nd = f_host.shape
    ofs, _, _ = O.radon_struct((nd, nd), angles, n_detectors=nd)
    kw = dict(alpha=ALPHA, mu=MU, L=L, coupling="frobenius", discrepancy="KL", dim=3, lib=lib)
    t0 = time.perf_counter()
    cref.solve(f_host, ofs, maxiter=1, **kw)
    t1 = time.perf_counter()
    cref.solve(f_host, ofs, maxiter=1 + iters, **kw)
    t2 = time.perf_counter()
    per_iter = ((t2 - t1) - (t1 - t0)) / iters
    return per_iter if per_iter > 0 else (t2 - t1) / (1 + iters)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(f_host, angles, L, iters, single_iters=1):
    """Time the FP32 C/OpenMP restatement on this host (bounded samples): on
    all cores of this process, and on one thread in a child process."""
    lib_path, arch = _native_cref()
    per_iter = _time_cref(lib_path, f_host, angles, L, iters)
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    C, Nz, na, nd = f_host.shape
    vox = C * Nz * nd * nd
    one = None
    if single_iters > 0:
        tmp = tempfile.mkdtemp(prefix="cref1_")
        fpath = os.path.join(tmp, "f.npy")
        import numpy as np
        np.save(fpath, f_host)
        np.save(os.path.join(tmp, "angles.npy"), np.asarray(angles, np.float64))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        code = (f"import sys, numpy as np; sys.path.insert(0, {ROOT!r}); import bench; "
                f"f = np.load({fpath!r}); a = np.load({os.path.join(tmp, 'angles.npy')!r}); "
                f"print(bench._time_cref({lib_path!r}, f, a, {float(L)!r}, {int(single_iters)}))")
        try:
            out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                                 text=True, timeout=600, check=True)
            one = vox / float(out.stdout.strip().splitlines()[-1])
        except Exception as exc:  # the 1-thread figure is informative only
            print(f"# 1-thread baseline failed: {exc}", file=sys.stderr, flush=True)
    return dict(value=vox / per_iter, unit="voxel-iters/s", cores=threads, kind="port",
                cpu=_cpu_model(), value_1thread=one,
                sample=f"same {nd}x{nd}x{Nz}x{C} volume, {iters} iterations on {threads} threads "
                       f"and {single_iters} on 1 thread, each timed as the difference of an "
                       f"(n+1)- and a 1-iteration run (allocation excluded), gcc -O3 {arch} "
                       f"OpenMP, FP32 restatement of the reference kernels",
                seconds_per_iter=per_iter)


def batched_main(args):
    """Config: 64 independent 2D multi-channel TGV-KL problems of uneven size,
    angle and channel counts advanced by one batch plan (one launch per kernel
    class and phase, hipGraph replay), tiles from the autotuner."""
    import numpy as np
    import torch

    from graptor_amd.autotune import autotune
    from graptor_amd.batch import BatchPlan
    from graptor_amd.phantoms import batched_2d_problem_set
    from graptor_amd.solver import TgvProblem

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    specs = batched_2d_problem_set(n=args.nprob, device=dev)
    probs = []
    for sp in specs:
        cpl = "frobenius" if sp["C"] > 1 else "uncorrelated"
        mu = [0.02] + [0.5] * (sp["C"] - 1)
        probs.append(TgvProblem(sp["f"], sp["geometry"], ALPHA, mu[:sp["C"]], dim=2,
                                coupling=cpl, discrepancy="KL", normest_mode="power",
                                device=dev, fused=False))
    plan = BatchPlan(probs)
    tune = autotune(probs, iters=3, plan=plan)
    print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    plan.iterate_graph(args.warmup, block=10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.iterate_graph(args.steps, block=10)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    vox = plan.voxels()
    value = vox * args.steps / wall
    ms = plan.profile(5)
    # accuracy vs the FP32 restatement of the reference kernels on a subset
    acc = None
    if not args.no_cpu_baseline:
        from oracle import cref
        from oracle import tgv_oracle as O
        errs = []
        sub = sorted(range(len(probs)), key=lambda i: probs[i].voxels())[:args.check]
        chk = [TgvProblem(specs[i]["f"], specs[i]["geometry"], ALPHA, probs[i].mu, dim=2,
                          coupling=probs[i].coupling, discrepancy="KL", L=probs[i].k.L,
                          device=dev) for i in sub]
        BatchPlan(chk).iterate(args.check_iters)
        torch.cuda.synchronize()
        for i, q in zip(sub, chk):
            ofs, _, _ = O.radon_struct((q.nx, q.ny), q.geometry.angles, n_detectors=q.nd)
            r = cref.solve(q.f.cpu().numpy(), ofs, ALPHA, q.mu, args.check_iters, q.k.L,
                           q.coupling, "KL", dim=2)
            gu = q.u.cpu().numpy()
            errs.append(float(np.linalg.norm(gu - r["u"]) / max(np.linalg.norm(r["u"]), 1e-30)))
        acc = {"problems": len(sub), "iterations": args.check_iters,
               "max_rel_l2_vs_fp32_restatement": max(errs), "per_problem": errs}
    line = {
        "metric": "voxel-iters/s (batched 2D coupled TGV-KL)",
        "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded multi-channel ellipse slices, Poisson counts)",
        "config": {"workload": f"{len(probs)} independent 2D TGV-KL problems, N in 96..384, "
                               "30..180 angles in [-70,70) deg, 1..4 channels, one batch plan, "
                               "hipGraph of 10 iterations",
                   "voxel_channels": vox, "tiles": {"fwd_band": tune["fwd_band"],
                                                    "adj_tile_y": tune["adj_tile_y"]}},
        "phases_ms": {"dual": ms[0] / 5, "fwd_prox": ms[1] / 5, "adj": ms[2] / 5,
                      "primal": ms[3] / 5},
        "accuracy": acc,
    }
    print(json.dumps(line), flush=True)


def warm_start_summary(r):
    """The warm-start object of a bench line from warmstart.evaluate()."""
    t = r["targets"]
    best = next((k for k in ("0.001", "0.003", "0.01") if t[k]["time_ratio_all"]), None)
    return {
        "value": t["0.001"]["time_ratio_all"],
        "unit": "x (plain PD time / (warm-start inference + PD) time until every slice is "
                "within 1e-3 relative L2 of u*)",
        "tightest_tolerance_both_reach": best,
        "value_at_tightest": t[best]["time_ratio_all"] if best else None,
        "slices": r["slices"], "seeds": r["seeds"],
        "targets": t, "error_median": r["error_median"],
        "ustar": {"iterations": r["ustar_iters"],
                  "largest_rel_change_last_1000": r["ustar_last1000_rel_change_max"],
                  "largest_rel_distance_from_iteration": r.get("ustar_distance_to_earlier")},
        "inference_ms": r["inference_ms"], "iteration_ms_batch": r["iteration_ms_batch"],
        "check_every": r["check_every"], "max_iters": r["max_iters"],
    }


def warmstart_main(args):
    """Config: learned warm start.  A 4-level U-Net (base 32, ~2M parameters,
    BF16 MFMA kernels, FP32 accumulation) trained on the solver's
    reconstructions of seeded 4-channel 256 x 256 ellipse slices (60 tilts in
    [-75, 75] deg, 256 bins, Poisson counts scaled as the reference's caller
    does, Frobenius-coupled TGV-KL).  --ws-stage train: make the training
    pairs with the GPU solver, train, save the weights (--save-net).
    --ws-stage eval: load the weights (--weights, default the committed ones)
    and compare warm start + PD with plain PD on unseen slices against u*
    (the plain iteration run until its change over 1000 iterations is below
    --ustar-tol): iterations and HIP-event time to 1e-2, 3e-3 and 1e-3."""
    import torch

    from graptor_amd import warmstart as ws

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log = lambda msg: print(f"# {msg}", file=sys.stderr, flush=True)
    if args.ws_stage == "train":
        net, info = ws.train_stage(n_train=args.train_pairs, target_iters=args.target_iters,
                                   train_steps=args.train_steps, device=dev, log=log)
        path = args.save_net or ws.WEIGHTS_PATH
        ws.save_weights(net, path, meta=info)
        print(json.dumps({"stage": "train", "weights": path, **info}), flush=True)
        return
    seeds = range(args.ws_seed0, args.ws_seed0 + args.eval_slices)
    if args.ws_stage == "ustar":
        r = ws.ustar_check(seeds=seeds, ustar_iters=args.ustar_cap,
                           ratios=[float(x) for x in args.step_ratios.split(",") if x],
                           iters=args.ratio_iters, device=dev, log=log)
        print(json.dumps({"stage": "ustar", **r}, default=float), flush=True)
        return
    net = ws.load_weights(args.weights or ws.WEIGHTS_PATH, device=dev)
    r = ws.evaluate(net, seeds=seeds, check_every=args.check_every, max_iters=args.max_iters,
                    ustar_min=args.ustar_iters, ustar_tol=args.ustar_tol,
                    ustar_cap=args.ustar_cap, device=dev, log=log)
    summ = warm_start_summary(r)
    line = {
        "metric": "time-to-1e-3 rel err: warm-start vs plain PD",
        "value": summ["value"], "unit": summ["unit"],
        "n_gpus": 1, "steps": args.max_iters, "warmup": 0,
        "ms_per_step": r["iteration_ms_batch"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 solver, bf16 network (fp32 accumulate)",
        "data": "synthetic (seeded 4-channel ellipse slices, Poisson counts, percentile-scaled)",
        "config": {"workload": f"warm start on {args.eval_slices} unseen 256x256x4 slices "
                               f"(seeds {seeds[0]}..{seeds[-1]}), 60 tilts, Frobenius TGV-KL; "
                               f"u* until change <= {args.ustar_tol:g} per 1000 it "
                               f"(>= {args.ustar_iters} it)",
                   "weights": os.path.relpath(args.weights or ws.WEIGHTS_PATH, ROOT),
                   "weights_meta": getattr(net, "meta", {})},
        "warm_start": summ,
    }
    print(json.dumps(line, default=float), flush=True)


def ellipse_main(args):
    """Config: plain 2D single-channel TGV-L2 (128x128 ellipse phantom in a 183
    grid, 90 angles in [-70, 70), 1% Gaussian noise, alpha (4, 1), mu 1, 500
    iterations) on the GPU and, for the same data and L, on this host's CPU
    with the FP32 C restatement of the reference kernels (all threads and one
    thread); the relative L2 difference of the two results is reported."""
    import numpy as np
    import torch

    from graptor_amd.phantoms import ellipse_problem_2d
    from graptor_amd.solver import TgvProblem
    from oracle import cref
    from oracle import tgv_oracle as O

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    iters = 500
    f, angles, ph, geom = ellipse_problem_2d(device=dev)
    ofs = geom.ofs.T.copy()
    L = O.normest_power(ofs, 183, 183, 183)
    prob = TgvProblem(f, geom, (4.0, 1.0), [1.0], dim=2, coupling="uncorrelated",
                      discrepancy="L2", L=L, device=dev)
    fh = prob.f.cpu().numpy()
    warm = prob.clone()
    warm.iterate(50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prob.iterate(iters)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    u = prob.u.cpu().numpy()
    vox = prob.voxels()
    # the same iterations replayed from a captured hipGraph (one-problem plan)
    from graptor_amd.batch import BatchPlan
    gprob = TgvProblem(f, geom, (4.0, 1.0), [1.0], dim=2, coupling="uncorrelated",
                       discrepancy="L2", L=L, device=dev)
    plan = BatchPlan([gprob])
    plan.iterate_graph(50, block=50)
    torch.cuda.synchronize()
    gprob2 = TgvProblem(f, geom, (4.0, 1.0), [1.0], dim=2, coupling="uncorrelated",
                        discrepancy="L2", L=L, device=dev)
    plan2 = BatchPlan([gprob2])
    plan2.iterate_graph(50, block=50)   # capture, then restart the state
    for t in (gprob2.u, gprob2.ubar, gprob2.p, gprob2.pbar, gprob2.v, gprob2.w, gprob2.lam):
        t.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan2.iterate_graph(iters, block=50)
    torch.cuda.synchronize()
    graph_s = time.perf_counter() - t0
    graph_err = float(np.linalg.norm(gprob2.u.cpu().numpy() - u) / np.linalg.norm(u))
    del plan, plan2
    lib, arch, path = None, "x86-64-v3", ""
    try:
        tmp = tempfile.mkdtemp(prefix="cref_")
        path = os.path.join(tmp, "libcpuref_native.so")
        subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c11",
                        os.path.join(ROOT, "oracle", "cpu_ref.c"), "-o", path, "-lm"],
                       check=True, capture_output=True, timeout=120)
        lib, arch = cref.load(path), "native"
    except Exception:
        lib = cref.load()
    kw = dict(alpha=(4.0, 1.0), mu=[1.0], L=L, coupling="uncorrelated", discrepancy="L2",
              dim=2, lib=lib)
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    r = cref.solve(fh, ofs, maxiter=iters, **kw)
    cpu_s = time.perf_counter() - t0
    env1 = dict(os.environ, OMP_NUM_THREADS="1")
    code = ("import sys,time,numpy as np; sys.path.insert(0, %r); from oracle import cref; "
            "d=np.load(%r); lib=cref.load(%r) if %r else cref.load(); t=time.perf_counter(); "
            "cref.solve(d['f'], d['ofs'], alpha=(4.0,1.0), mu=[1.0], maxiter=%d, L=%r, "
            "coupling='uncorrelated', discrepancy='L2', dim=2, lib=lib); "
            "print(time.perf_counter()-t)")
    npz = os.path.join(tempfile.mkdtemp(prefix="c1_"), "c1.npz")
    np.savez(npz, f=fh, ofs=ofs)
    one = subprocess.run([sys.executable, "-c", code % (ROOT, npz, path if arch == "native"
                                                        else "", arch == "native", iters, L)],
                         env=env1, capture_output=True, text=True, timeout=600)
    cpu1_s = float(one.stdout.strip().splitlines()[-1]) if one.returncode == 0 else None
    err = float(np.linalg.norm(u - r["u"]) / np.linalg.norm(r["u"]))
    line = {
        "metric": "voxel-iters/s (2D TGV-L2, 183x183, 90 angles)",
        "value": vox * iters / gpu_s, "unit": "voxel-iters/s", "n_gpus": 1, "steps": iters,
        "warmup": 50, "ms_per_step": gpu_s * 1e3 / iters, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipse phantom, 1% Gaussian noise)",
        "config": {"workload": "2D single-channel TGV-L2, 128x128 ellipses in a 183 grid, "
                               "90 angles in [-70,70), alpha (4,1), mu 1, 500 iterations"},
        "cpu_baseline": {"value": vox * iters / cpu_s, "unit": "voxel-iters/s", "cores": threads,
                         "kind": "port", "sample": f"the same 500 iterations, gcc -O3 {arch} "
                                                   "OpenMP",
                         "one_thread_value": (vox * iters / cpu1_s) if cpu1_s else None},
        "rel_l2_gpu_vs_cpu": err,
        "graph_replay": {"value": vox * iters / graph_s, "ms_per_step": graph_s * 1e3 / iters,
                         "rel_l2_vs_eager": graph_err},
    }
    print(json.dumps(line), flush=True)


def unet_main(args):
    """The warm-start network's kernels alone: Adam steps of the 4-level
    U-Net (base 32) on a batch of 16 random 256 x 256 x 4 inputs and forward
    passes of a batch of 32, timed with HIP events; MFMA roofline of the
    training step (3 x forward flops) and of inference."""
    import torch

    from graptor_amd.warmstart import UNet, conv_flops

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    net = UNet(cin=4, cout=4, base=32, device=dev, seed=0)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    X = torch.randn((16, 256, 256, 4), device=dev, generator=g)
    T = torch.randn((16, 256, 256, 4), device=dev, generator=g)
    Xi = torch.randn((32, 256, 256, 4), device=dev, generator=g)
    ev = lambda: torch.cuda.Event(enable_timing=True)

    def train_step():
        net.forward(X)
        net.backward(T, return_loss=False)
        net.adam()

    for _ in range(max(args.warmup, 1)):
        train_step()
        net.forward(Xi)
    steps = args.steps
    a, b = ev(), ev()
    a.record()
    for _ in range(steps):
        train_step()
    b.record()
    b.synchronize()
    train_ms = a.elapsed_time(b) / steps
    a.record()
    for _ in range(steps):
        net.forward(Xi)
    b.record()
    b.synchronize()
    infer_ms = a.elapsed_time(b) / steps
    f = conv_flops(net, 256, 256)
    train_tf = 3 * f * 16 / (train_ms * 1e-3) / 1e12
    infer_tf = f * 32 / (infer_ms * 1e-3) / 1e12
    line = {
        "metric": "U-Net training / inference throughput (warm-start network)",
        "value": 16 / (train_ms * 1e-3), "unit": "training images/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": train_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)", "data": "synthetic (random)",
        "config": {"workload": "4-level U-Net, base 32, 1.95 M parameters, batch 16 "
                               "(training) / 32 (inference), 256x256x4"},
        "roofline": {"bound": "mfma", "achieved": infer_tf, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": infer_tf / 2500.0, "traffic": None,
                     "kernel": "forward pass (conv_tile_kernel and the elementwise kernels)"},
        "train_tflops": train_tf, "inference_ms_batch32": infer_ms,
    }
    print(json.dumps(line), flush=True)


def slab_main(args):
    """One GPU's share of the 512x512x256x4, 90-tilt volume split into 8
    z-slabs: a 512x512x32x4 slab with halo planes on both sides (a middle
    rank), iterated with the halo-aware kernels, and the packing of its two
    halo exchanges per iteration (the buffers a rank sends to its neighbours
    over xGMI)."""
    import torch

    from graptor_amd.phantoms import coupled_granule_problem
    from graptor_amd.solver import TgvProblem

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N, nz, na, ranks = 512, 32, 90, 8
    f, angles, phantom, geom = coupled_granule_problem(N=N, Nz=nz, na=na, device=dev)
    del phantom
    prob = TgvProblem(f, geom, ALPHA, MU, dim=3, coupling="frobenius", discrepancy="KL",
                      normest_mode="power", device=dev)
    C = prob.C
    plane = (C, N, N)
    halo = {"ubar_hi": torch.empty(plane, device=dev), "pbar_lo": torch.empty((3,) + plane, device=dev),
            "v_lo": torch.empty(plane, device=dev), "w_hi": torch.empty((3,) + plane, device=dev)}
    # neighbours' planes stand in as copies of the slab's own edge planes
    def exchange():
        halo["ubar_hi"].copy_(prob.ubar[:, -1])
        halo["pbar_lo"].copy_(prob.pbar[:, :, 0])
        halo["v_lo"].copy_(prob.v[2, :, 0])
        for j, m in enumerate((4, 5, 2)):
            halo["w_hi"][j].copy_(prob.w[m, :, -1])
    exchange()
    prob.state.halo_ubar_hi = halo["ubar_hi"].data_ptr()
    prob.state.halo_pbar_lo = halo["pbar_lo"].data_ptr()
    prob.state.halo_v_lo = halo["v_lo"].data_ptr()
    prob.state.halo_w_hi = halo["w_hi"].data_ptr()
    prob.iterate(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prob.iterate(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # what the rank sends per iteration: its own edge planes packed into one
    # buffer per neighbour and phase (gr_tgv_pack_halos), timed on its stream
    import ctypes

    from graptor_amd import _lib
    lib = _lib.load()
    send = {"dual_below": torch.empty(plane, device=dev),
            "dual_above": torch.empty((3,) + plane, device=dev),
            "primal_below": torch.empty((3,) + plane, device=dev),
            "primal_above": torch.empty(plane, device=dev)}
    st = prob._stream()

    def pack():
        for ph in ("dual", "primal"):
            _lib.check(lib.gr_tgv_pack_halos(
                ctypes.byref(prob.params), ctypes.byref(prob.state), _lib.HALO_PHASE[ph],
                send[ph + "_below"].data_ptr(), send[ph + "_above"].data_ptr(), st), "pack_halos")
    pack()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        pack()
    e1.record(stream)
    e1.synchronize()
    pack_ms = e0.elapsed_time(e1) / args.steps
    halo_bytes = sum(t.numel() for t in halo.values()) * 4  # received per iteration
    line = {
        "metric": "voxel-iters/s (one z-slab of a slab-sharded coupled TGV-KL volume)",
        "value": prob.voxels() * args.steps / wall, "unit": "voxel-iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded granule phantom, Poisson counts)",
        "config": {"workload": "512x512x32 slab x 4 channels (1/8 of 512x512x256), 90 tilts, "
                               "halo planes above and below", "ranks_modelled": ranks},
        "halo": {"bytes_received_per_iteration": halo_bytes,
                 "pack_ms_per_iteration": pack_ms,
                 "xgmi_ms_per_iteration_at_150GBps_per_link": halo_bytes / 2 / 150e9 * 1e3},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed iterations")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-iters", type=int, default=50)
    ap.add_argument("--ramp-iters", type=int, default=200,
                    help="untimed set-up iterations ahead of the profiled pass and warm-up")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--workload", default="coupled3d",
                    choices=["coupled3d", "large3d", "batched2d", "warmstart", "unet",
                             "ellipse2d", "slab"])
    ap.add_argument("--cap-iters", type=int, default=2000,
                    help="large3d: iteration cap of the run to a relative-change tolerance")
    ap.add_argument("--change-tol", type=float, default=1e-3)
    ap.add_argument("--train-pairs", type=int, default=20000)
    ap.add_argument("--train-steps", type=int, default=12000)
    ap.add_argument("--eval-slices", type=int, default=256)
    ap.add_argument("--target-iters", type=int, default=2500)
    ap.add_argument("--ustar-iters", type=int, default=100000,
                    help="least iterations of the u* run")
    ap.add_argument("--budget-iters", type=int, default=2500)
    ap.add_argument("--max-iters", type=int, default=60000,
                    help="iteration cap of the plain and warm-started arms")
    ap.add_argument("--save-net", default=None, help="write the trained weights (.npz)")
    ap.add_argument("--weights", default=None, help="warm-start weights (.npz)")
    ap.add_argument("--ws-stage", default="eval", choices=["train", "eval", "ustar"],
                    help="ustar: distance of a --ustar-cap-iteration u* to runs with other "
                         "primal/dual step ratios (same fixed point)")
    ap.add_argument("--step-ratios", default="0.25,0.5,2,4")
    ap.add_argument("--ratio-iters", type=int, default=120000)
    ap.add_argument("--ws-seed0", type=int, default=60000)
    ap.add_argument("--ws-slices", type=int, default=16,
                    help="slices of the warm-start comparison in the default line")
    ap.add_argument("--ustar-tol", type=float, default=1e-5)
    ap.add_argument("--ustar-cap", type=int, default=200000)
    ap.add_argument("--check-every", type=int, default=250)
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--nprob", type=int, default=64)
    ap.add_argument("--check", type=int, default=8)
    ap.add_argument("--check-iters", type=int, default=50)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (known bytes) so that "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE can be scaled")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU transport "
              "(one rank's z-slab with halo planes: --workload slab); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched2d":
        return batched_main(args)
    if args.workload == "warmstart":
        return warmstart_main(args)
    if args.workload == "unet":
        return unet_main(args)
    if args.workload == "ellipse2d":
        return ellipse_main(args)
    if args.workload == "slab":
        return slab_main(args)

    import numpy as np
    import torch

    from graptor_amd.phantoms import coupled_granule_problem
    from graptor_amd.solver import TgvProblem

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    large = args.workload == "large3d"
    N, nz, na = (512, 256, 90) if large else (256, args.nz, 60)
    f, angles, phantom, geom = coupled_granule_problem(N=N, Nz=nz, na=na, device=dev)
    del phantom
    prob = TgvProblem(f, geom, ALPHA, MU, dim=3, coupling="frobenius", discrepancy="KL",
                      normest_mode="power", device=dev)
    del f
    print(f"# problem ready: {prob.C}x{prob.Nz}x{prob.ny}x{prob.nx}, na={prob.na}, "
          f"L={prob.k.L:.6f}", file=sys.stderr, flush=True)
    # set-up ahead of the W warm-up steps: a fixed clock ramp (a 1000-iteration
    # solve runs at the clocks of a busy GPU, the first few of a cold one do
    # not), then the per-kernel durations, HIP events around every launch of a
    # profiled pass (one-stream order)
    prob.iterate(args.ramp_iters)
    ms = prob.profile(args.profile_iters)
    prob.iterate(args.warmup)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    prob.iterate(args.steps)
    end.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    gpu_ms = start.elapsed_time(end)
    value = prob.voxels() * args.steps / wall
    print(f"# timed {args.steps} iterations: wall {wall:.4f}s, events {gpu_ms / 1e3:.4f}s",
          file=sys.stderr, flush=True)

    counts = launch_counts(args.profile_iters, prob.u2 is not None)
    nbytes = kernel_bytes(prob)
    kern = {k: dict(avg_ms=m / c, launches=c, bytes=nbytes[k],
                    gbs=nbytes[k] / (m / c * 1e-3) / 1e9)
            for k, m, c in zip(KERNELS, ms, counts) if c > 0}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    achieved = kern[dom]["gbs"]
    print("# kernels: " + json.dumps(kern), file=sys.stderr, flush=True)

    if args.copy_calibration:
        src = torch.ones(256 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("# copy calibration: 3 copies of 1 GiB (read 1 GiB, write 1 GiB each)",
              file=sys.stderr, flush=True)
        del src, dst

    cpu = None
    # the C restatement keeps the reference's padded layouts (~0.4 KB per
    # voxel-channel): the large volume would need ~100 GB of host memory
    if not args.no_cpu_baseline and not large:
        f_host = prob.f.cpu().numpy()
        cpu = cpu_baseline(f_host, geom.angles, prob.k.L, args.cpu_iters)
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)
        cpu.pop("seconds_per_iter", None)

    warm = None
    if not args.no_warm_start and not large:
        # the second half of the metric on a bounded sample: committed weights,
        # --ws-slices unseen slices
        from graptor_amd import warmstart as ws
        wpath = args.weights or ws.WEIGHTS_PATH
        if os.path.exists(wpath):
            net = ws.load_weights(wpath, device=dev)
            r = ws.evaluate(net, seeds=range(args.ws_seed0, args.ws_seed0 + args.ws_slices),
                            check_every=args.check_every, max_iters=args.max_iters,
                            ustar_min=args.ustar_iters, ustar_tol=args.ustar_tol,
                            ustar_cap=args.ustar_cap, device=dev,
                            log=lambda m: print(f"# {m}", file=sys.stderr, flush=True))
            warm = warm_start_summary(r)
            warm["weights"] = os.path.relpath(wpath, ROOT)
            del net
        else:
            print(f"# no warm-start weights at {wpath}", file=sys.stderr, flush=True)

    converge = None
    if large and args.cap_iters > 0:
        # run from zero to the iteration cap, stopping when the relative change
        # of u over 50 iterations drops to change_tol (checks not timed)
        run = TgvProblem(prob.f, geom, ALPHA, MU, dim=3, coupling="frobenius",
                         discrepancy="KL", L=prob.k.L, feasible=False, device=dev)
        del prob.u, prob.ubar, prob.p, prob.pbar, prob.v, prob.w, prob.ktl
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        it, ms, change, prev = 0, 0.0, None, run.u.clone()
        while it < args.cap_iters:
            e0.record()
            run.iterate(50)
            e1.record()
            e1.synchronize()
            ms += e0.elapsed_time(e1)
            it += 50
            change = float((run.u - prev).norm() / run.u.norm().clamp(min=1e-30))
            prev.copy_(run.u)
            if change <= args.change_tol:
                break
        converge = {"iterations": it, "seconds": ms / 1e3, "rel_change_per_50": change,
                    "tolerance": args.change_tol, "cap": args.cap_iters,
                    "memory_GiB": torch.cuda.max_memory_allocated(dev) / 2 ** 30}
        print(f"# converge: {converge}", file=sys.stderr, flush=True)

    counted = {}
    tpath = os.path.join(ROOT, "profiles", "coupled3d_traffic.json")
    if not large and args.nz == 64 and os.path.exists(tpath):
        # HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
        # (scaled by a 1 GiB copy) for this workload, one-stream launches
        with open(tpath) as fh:
            table = json.load(fh)["kernels_MiB_per_launch"]
        for k, key in (("dual_single_kernel", "dual_single_kernel"),
                       ("fwd_single_kernel", "fwd_single_kernel_profiled"),
                       ("adj_single_kernel", "adj_single_kernel"),
                       ("primal_single_kernel", "primal_single_kernel")):
            if key in table:
                counted[k] = (table[key]["fetch"] + table[key]["write"]) * 2 ** 20
    traffic = counted.get(dom)
    # the projectors, by the same algorithmic-bytes rule: bound by vector ALU
    # and LDS work, so their byte rate is far below the HBM peak; their vector
    # rate counts the useful multiply-adds, two per pixel, angle and slice
    # (each pixel meets two bins per angle), against the non-packed FP32 peak
    fma = 2.0 * prob.nx * prob.ny * prob.na * prob.C * prob.Nz
    projector = {}
    for k in ("fwd_single_kernel", "adj_single_kernel"):
        if k not in kern:
            continue
        tflops = 2.0 * fma / (kern[k]["avg_ms"] * 1e-3) / 1e12
        projector[k] = {"avg_launch_ms": kern[k]["avg_ms"], "bytes_per_launch": kern[k]["bytes"],
                        "achieved": kern[k]["gbs"], "unit": "GB/s",
                        "frac": kern[k]["gbs"] / HBM_PEAK_GBS, "traffic": counted.get(k),
                        "valu": {"useful_fma_per_launch": fma, "achieved": tflops,
                                 "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": tflops / VALU_PEAK_TFLOPS}}

    line = {
        "metric": "voxel-iters/s (coupled TGV-KL) + time-to-1e-3 rel err: warm-start vs plain PD",
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded granule phantom, Poisson counts)",
        "config": {"workload": "3D Frobenius-coupled 4-channel TGV-KL, %dx%dx%d slices, "
                               "%d tilts in [-75,75] deg, %d bins" % (N, N, nz, na, N),
                   "volume": [4, nz, N, N], "angles": na, "bins": N,
                   "alpha": list(ALPHA), "mu": list(MU), "iterations_timed": args.steps},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "bytes_per_launch": kern[dom]["bytes"],
                     "avg_launch_ms": kern[dom]["avg_ms"]},
        "kernels_ms": {k: v["avg_ms"] for k, v in kern.items()},
        "projector": dict(bound="valu+lds", **projector),
        "cpu_baseline": cpu,
    }
    if warm is not None:
        line["warm_start"] = warm
    if converge is not None:
        line["convergence"] = converge
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
