#!/usr/bin/env python3
"""Benchmark of the 3D VTI time step on one MI355X.

Workload: the 3D VTI forward + RTM configuration -- a 400x512x512 interior
with a 30-point boundary, i.e. a 460x572x572 padded grid, 12.5 m spacing,
dt = 1 ms, seeded layered model with dipping interfaces (v 1500-4500 m/s,
eps 0-0.25, delta 0-0.12).  A step is one forward time level of the coupled
sigma_x/sigma_z system with absorbing damping (the reference's
forwardKernel + injectSource + abcXYZ/abcXY over all z-blocks).

Prints one JSON line: Gpts/s over the padded grid, the roofline of the step
kernel (HIP-event timing on the propagator's stream; algorithmic bytes 36 per
point), and the CPU baseline (the oracle's C/OpenMP restatement of the same
step, timed here on a bounded number of steps of the same grid).  The
stencil arithmetic is the reference's own (double weights and sums over FP32
storage, kernel3d.cu:54-69); "extra" also reports the FP32-arithmetic
variant and the fused backward (gradient, RTM) steps on the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "thesis-source-code_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FWD = 36  # sigma_x, sigma_z cur+prev, c11, c13, c33 read; sigma_x, sigma_z written
BYTES_BWD = 84  # fused gradient step: 14 reads + 7 writes of 4 bytes per point
BYTES_RTM = 68  # fused RTM step: 12 reads + 5 writes


def layered(nz, ny, nx, seed=2024):
    from vtiamd import workloads as W

    return W.layered3d(nz, ny, nx, seed)


def progress(msg: str) -> None:
    """One line on stderr as each part of the run ends (the JSON result line
    goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def load_traffic(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(f"{kernel}@{workload}")
    except (OSError, ValueError):
        return None


def cpu_baseline(ved, npad, dx, dt, budget_s=6.0):
    """The reference's CPU path restated in C/OpenMP (the oracle's cpu3d.c and
    its 2D loop restatement), timed in this process on the host threads this
    job is given (OMP_NUM_THREADS; the box gives one GPU's job 16), on the
    same grids as the GPU legs:
    - 3D forward step (laplacian3d.ispc forward + the modeling loop and abc
      of vtifd3d8o-modeling.cpp) on the full 460x572x572 grid, >= 100 steps:
      the headline CPU figure;
    - 3D gradient step (gradientCij + forward on the adjoint + abc,
      vtifd3d8o-objfuncGrad.cpp:322-367), a window on the full grid;
    - 2D cfg1 modeling (synthetic_f, 256x256 + 40, 2000 steps) and cfg2 shot
      gradient (objFuncGradient_f, 512x1024 + 40, 3000 steps), both in full;
    - cfg4: the per-shot gradient at the mean aperture (a 500-step window,
      scaled to 2500 steps) times 256 shots."""
    import oracle
    from vtiamd import workloads as W

    v, e, d = ved
    nz, ny, nx = v.shape
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    legs = {}
    rng = np.random.default_rng(0)
    st = [np.zeros((nz, ny, nx), np.float32) for _ in range(4)]
    st[1][:] = rng.standard_normal((nz, ny, nx), dtype=np.float32) * 1e-3
    oracle.cpu3d_steps(0, st, v, e, d, npad, dx, dx, dx, dt, 1)  # first touch
    n = 100
    t0 = time.perf_counter()
    oracle.cpu3d_steps(0, st, v, e, d, npad, dx, dx, dx, dt, n)
    el = time.perf_counter() - t0
    gpts = nz * ny * nx * n / el / 1e9
    legs["3d_forward_step"] = {"gpts_s": round(gpts, 4), "steps": n, "s": round(el, 2)}
    progress(f"cpu 3d forward leg: {el:.1f} s")
    lam = [np.zeros((nz, ny, nx), np.float32) for _ in range(4)]
    lam[1][:] = rng.standard_normal((nz, ny, nx), dtype=np.float32) * 1e-6
    gc = [np.zeros((nz, ny, nx), np.float32) for _ in range(3)]
    oracle.cpu3d_backward_steps(st, lam, gc, v, e, d, npad, dx, dx, dx, dt, 1)
    t0 = time.perf_counter()
    nb = 20
    oracle.cpu3d_backward_steps(st, lam, gc, v, e, d, npad, dx, dx, dx, dt, nb)
    elb = time.perf_counter() - t0
    legs["3d_gradient_step"] = {"gpts_s": round(nz * ny * nx * nb / elb / 1e9, 4), "steps": nb, "s": round(elb, 2)}
    progress(f"cpu 3d gradient leg: {elb:.1f} s")
    del st, lam, gc

    # 2D cfg1: synthetic_f in full
    nx1, nz1, np1, nt1 = 256, 256, 40, 2001
    nnx1, nnz1 = nx1 + 2 * np1, nz1 + 2 * np1
    v1 = 1500.0 + 2500.0 * np.arange(nnz1, dtype=np.float32)[:, None] / nnz1 * np.ones((1, nnx1), np.float32)
    c1 = (v1 * v1).astype(np.float32)
    taper1 = _taper2d(nx1, nz1, np1)
    w1 = W.ricker(nt1, 1e-3, 15.0, 0.1)
    r1 = np.array([[np1 + i, np1 + 1] for i in range(nx1)], np.int32)
    s1 = np.array([[np1 + 128, np1 + 20, nx1, 0]], np.int32)
    t0 = time.perf_counter()
    oracle.synthetic2d(oracle.F32, c1 * np.float32(1.2), c1 * np.float32(1.05), c1, w1, s1, r1, taper1, nx1, nz1,
                       nt1, np1, 10.0, 10.0, 1e-3, 4e-3)
    el1 = time.perf_counter() - t0
    legs["2d_cfg1_synthetic_full"] = {"s": round(el1, 2), "shots_per_s": round(1 / el1, 3),
                                      "gpts_s": round(nnx1 * nnz1 * (nt1 - 1) / el1 / 1e9, 4)}
    # 2D cfg2: one shot gradient in full
    P = W.cfg2()
    ci = W.cij(*P["init"])
    taper2 = _taper2d(P["nx"], P["nz"], P["npad"])
    sl = P["sloc"][16:17]
    nnt = (P["nt"] - 1) // 4 + 1
    obs = np.zeros((nnt, P["nx"]), np.float32)
    t0 = time.perf_counter()
    oracle.fg2d(oracle.F32, False, obs, *ci, None, P["w"], sl, P["rloc"], taper2, P["nx"], P["nz"], P["nt"],
                P["npad"], P["dx"], P["dz"], P["dt"], P["rate"])
    el2 = time.perf_counter() - t0
    progress(f"cpu 2d legs: {el1:.1f} s, {el2:.1f} s")
    legs["2d_cfg2_shot_gradient_full"] = {"s": round(el2, 2), "shot_gradients_per_s": round(1 / el2, 4),
                                          "call_of_32_shots_s": round(32 * el2, 1)}
    # cfg4: one shot at the mean aperture (600 columns), 500-step window, x 256 shots
    nx4, nt4w = 600, 501
    c4 = [np.ascontiguousarray(a[:, :nx4 + 2 * P["npad"]]) for a in ci]
    taper4 = _taper2d(nx4, P["nz"], P["npad"])
    r4 = np.array([[P["npad"] + i, P["npad"] + 1] for i in range(nx4)], np.int32)
    s4 = np.array([[P["npad"] + nx4 // 2, P["npad"] + 2, nx4, 0]], np.int32)
    t0 = time.perf_counter()
    oracle.fg2d(oracle.F32, False, np.zeros(((nt4w - 1) // 4 + 1, nx4), np.float32), *c4, None, P["w"][:nt4w], s4,
                r4, taper4, nx4, P["nz"], nt4w, P["npad"], P["dx"], P["dz"], P["dt"], P["rate"])
    el4 = (time.perf_counter() - t0) * 2500 / (nt4w - 1)
    legs["cfg4_256_shots"] = {"per_shot_s": round(el4, 2), "s": round(256 * el4, 1),
                              "shot_gradients_per_s": round(1 / el4, 4),
                              "sample": "one 600-column aperture, 500-step window scaled to 2500 steps"}
    return {
        "value": round(gpts, 4),
        "unit": "Gpts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} absorbing forward steps of the full {nz}x{ny}x{nx} grid on {threads} threads "
        f"(the job's host thread share), {el:.1f} s: the reference CPU path (laplacian3d.ispc forward + "
        f"vtifd3d8o-modeling.cpp loop and abc) restated in C/OpenMP, FP32; other legs in 'legs'",
        "legs": legs,
    }


def _taper2d(nx, nz, npad):
    import vtiamd

    return vtiamd.init_taper2d(nx, nz, npad)


def batch_gradient_2d(vtiamd, ctx, ns=256, nplain=32, seed=7, variants=None, graph=True):
    """Shot-gradients/s of the batched 2D gradient (cfg4: `ns` shots, the
    autotuned launch shape) against the plain one-shot-per-call driver on
    every (ns/nplain)-th of the same shots: a 4096-wide seeded model (nz 512,
    40-point boundary), each shot on its own aperture of 300-900 columns,
    2500 steps (rate 4 ms), a receiver on every column."""
    rng = np.random.default_rng(seed)
    NX, nz, npad, nt, dx, dt, rate = 4096, 512, 40, 2501, 10.0, 1e-3, 4e-3
    NNX, nnz = NX + 2 * npad, nz + 2 * npad
    z = np.arange(nnz, dtype=np.float32)[:, None] * np.ones((1, NNX), np.float32)
    x = np.ones((nnz, 1), np.float32) * np.arange(NNX, dtype=np.float32)[None, :]
    v = 1500.0 + 3000.0 * z / nnz + 100.0 * np.sin(x / 97.0)
    c33 = (v * v).astype(np.float32)
    c11 = (c33 * 1.2).astype(np.float32)
    c13 = (c33 * 1.05).astype(np.float32)
    t = np.arange(nt) * dt - 0.12
    a = (np.pi * 10.0 * t) ** 2
    w = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    shots, rloc = [], []
    for _ in range(ns):
        nx = int(rng.integers(300, 901))
        c0 = int(rng.integers(0, NX - nx + 1))
        shots.append((c0, nx, npad + nx // 2, npad + 2, nx, len(rloc)))
        rloc += [(npad + i, npad + 1) for i in range(nx)]
    shots = np.array(shots, np.int32)
    rloc = np.array(rloc, np.int32)
    nnt = (nt - 1) // int(rate / dt + 0.5) + 1
    data = np.zeros(int(sum(nnt * s[4] for s in shots)), np.float32)
    tuned = vtiamd.tune_batch2d(ctx, shots, rloc, NX, nz, nt, npad, dx, dx, dt, rate)
    if variants is None:
        shape = {k: tuned[k] for k in ("levels", "tile_rows", "max_wave")}
        # FP32 snapshots (bit-exact with the one-shot driver), then S16
        # snapshots (16 bits, within 1e-4 of the FP64 oracle)
        variants = {"autotuned": shape, "autotuned_s16_snapshots": {**shape, "snap": "s16"}}
    timed = {}
    for name, opts in variants.items():
        t0 = time.perf_counter()
        vtiamd.objfunc_gradient2d_batch(ctx, data, c11, c13, c33, w, shots, rloc, NX, nz, nt, npad, dx, dx, dt, rate,
                                        **opts)
        timed[name] = time.perf_counter() - t0
    tb = timed[next(iter(timed))]
    # the same shots captured once (hipGraph of the whole launch sequence) and
    # replayed on new models, as the iterations of an inversion do
    shape = {k: tuned[k] for k in ("levels", "tile_rows", "max_wave")}
    plan = graph_plan_rate(vtiamd, ctx, data, w, shots, rloc, (NX, nz, nt, npad, dx, dx, dt, rate), (c11, c13, c33),
                           shape, direct_s=tb) if graph else None
    # the plain driver, one call per shot on the cropped model
    tp = 0.0
    sample = list(range(0, ns, max(1, ns // nplain)))[:nplain]
    for k in sample:
        c0, nx, ix, iz, nr, r0 = (int(q) for q in shots[k])
        wd = nx + 2 * npad
        taper = vtiamd.init_taper2d(nx, nz, npad)
        crop = [np.ascontiguousarray(c[:, c0:c0 + wd]) for c in (c11, c13, c33)]
        d = np.zeros((nnt, nr), np.float32)
        sl = np.array([[ix, iz, nr, 0]], np.int32)
        if k == sample[0]:  # first call warms the driver up
            vtiamd.objfunc_gradient2d(ctx, d, *crop, w, sl, rloc[r0:r0 + nr], taper, nx, nz, nt, npad, dx, dx, dt,
                                      rate)
        t0 = time.perf_counter()
        vtiamd.objfunc_gradient2d(ctx, d, *crop, w, sl, rloc[r0:r0 + nr], taper, nx, nz, nt, npad, dx, dx, dt, rate)
        tp += time.perf_counter() - t0
    return {"shots": ns, "model_padded": [nnz, NNX], "aperture_cols": [300, 900], "steps": nt - 1,
            "batched_shot_gradients_per_s": round(ns / tb, 3), "batched_s": round(tb, 3),
            "plain_shot_gradients_per_s": round(nplain / tp, 3), "plain_shots_timed": nplain,
            "speedup": round((ns / tb) / (nplain / tp), 2), "launch_shape": tuned,
            **({"graph_plan": plan,
                "graph_plan_speedup_over_plain": round(plan["shot_gradients_per_s"] / (nplain / tp), 2),
                "small_grid": batch_small_2d(vtiamd, ctx)} if graph else {}),
            **({"variants_shot_gradients_per_s": {k: round(ns / v, 3) for k, v in timed.items()},
                "variants_speedup_over_plain": {k: round((ns / v) / (nplain / tp), 2) for k, v in timed.items()}}
               if variants else {})}


def graph_plan_rate(vtiamd, ctx, data, w, shots, rloc, geo, model, shape, direct_s, replays=3):
    """Shot-gradients/s of a vtiamd.BatchPlan replayed `replays` times on
    models scaled by 1 + 1e-3 k (capture time reported apart)."""
    ns = len(shots)
    t0 = time.perf_counter()
    with vtiamd.BatchPlan(ctx, data, w, shots, rloc, *geo, **shape) as plan:
        t_cap = time.perf_counter() - t0
        times = []
        for k in range(replays):
            m = [(c * np.float32(1 + 1e-3 * k)).astype(np.float32) for c in model]
            t1 = time.perf_counter()
            plan.run(*m)
            times.append(time.perf_counter() - t1)
    tr = min(times)
    return {"shot_gradients_per_s": round(ns / tr, 3), "replay_s": round(tr, 4), "capture_s": round(t_cap, 3),
            "direct_call_s": round(direct_s, 4), "replays": replays}


def batch_small_2d(vtiamd, ctx, ns=32, seed=5):
    """Where launches are short: 32 shots of 64-160 columns on a 1024 x 96
    model (+20 boundary), 1000 steps -- direct batched calls against graph
    replays of the same shots, and the plain one-shot driver."""
    rng = np.random.default_rng(seed)
    NX, nz, npad, nt, dx, dt, rate = 1024, 96, 20, 1001, 10.0, 1e-3, 4e-3
    NNX, nnz = NX + 2 * npad, nz + 2 * npad
    v = np.broadcast_to(1500.0 + 2000.0 * np.arange(nnz, dtype=np.float32)[:, None] / nnz, (nnz, NNX))
    c33 = np.ascontiguousarray(v * v, dtype=np.float32)
    c11, c13 = (c33 * 1.2).astype(np.float32), (c33 * 1.05).astype(np.float32)
    t = np.arange(nt) * dt - 0.1
    a = (np.pi * 15.0 * t) ** 2
    w = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    shots, rloc = [], []
    for _ in range(ns):
        nx = int(rng.integers(64, 161))
        c0 = int(rng.integers(0, NX - nx + 1))
        shots.append((c0, nx, npad + nx // 2, npad + 2, nx, len(rloc)))
        rloc += [(npad + i, npad + 1) for i in range(nx)]
    shots, rloc = np.array(shots, np.int32), np.array(rloc, np.int32)
    nnt = (nt - 1) // int(rate / dt + 0.5) + 1
    data = np.zeros(int(sum(nnt * s[4] for s in shots)), np.float32)
    geo = (NX, nz, nt, npad, dx, dx, dt, rate)
    vtiamd.objfunc_gradient2d_batch(ctx, data, c11, c13, c33, w, shots, rloc, *geo)
    td = []
    for _ in range(3):
        t0 = time.perf_counter()
        vtiamd.objfunc_gradient2d_batch(ctx, data, c11, c13, c33, w, shots, rloc, *geo)
        td.append(time.perf_counter() - t0)
    plan = graph_plan_rate(vtiamd, ctx, data, w, shots, rloc, geo, (c11, c13, c33), {}, direct_s=min(td))
    tp = 0.0
    for k in range(4):
        c0, nx, ix, iz, nr, r0 = (int(q) for q in shots[k])
        wd = nx + 2 * npad
        taper = vtiamd.init_taper2d(nx, nz, npad)
        crop = [np.ascontiguousarray(c[:, c0:c0 + wd]) for c in (c11, c13, c33)]
        d = np.zeros((nnt, nr), np.float32)
        sl = np.array([[ix, iz, nr, 0]], np.int32)
        t0 = time.perf_counter()
        vtiamd.objfunc_gradient2d(ctx, d, *crop, w, sl, rloc[r0:r0 + nr], taper, nx, nz, nt, npad, dx, dx, dt, rate)
        if k:
            tp += time.perf_counter() - t0
    plain = 3 / tp
    return {"shots": ns, "model_padded": [nnz, NNX], "aperture_cols": [64, 160], "steps": nt - 1,
            "direct_shot_gradients_per_s": round(ns / min(td), 3), "graph_plan": plan,
            "plain_shot_gradients_per_s": round(plain, 3),
            "graph_plan_speedup_over_plain": round(plan["shot_gradients_per_s"] / plain, 2)}


def shot_gradient_2d(vtiamd, ctx, nshots=2, ns_call=32):
    """2D gradients of the cfg2 inversion (512x1024 + 40 boundary, 3000 steps,
    1024 receivers, rate 4 ms; initial model against data of the true model
    with its 64 anomalies): one shot per call (the one-shot loop), and one call
    with all 32 shots as the reference's configuration makes it (run as one
    batch inside objfunc_gradient2d): shot-gradients/s of both."""
    from vtiamd import workloads as W

    P = W.cfg2(nshots=ns_call)
    ct, ci = W.cij(*P["true"]), W.cij(*P["init"])
    nx, nz, npad, nt = P["nx"], P["nz"], P["npad"], P["nt"]
    taper = vtiamd.init_taper2d(nx, nz, npad)
    args = (taper, nx, nz, nt, npad, P["dx"], P["dz"], P["dt"], P["rate"])
    sl, rloc = P["sloc"], P["rloc"]
    nnt = (nt - 1) // 4 + 1
    # every shot records all 1024 receivers: rloc repeated per shot, sloc pointing at its block
    sl_all = sl.copy()
    sl_all[:, 3] = np.arange(ns_call) * nx
    rl_all = np.tile(rloc, (ns_call, 1))
    data = vtiamd.synthetic2d(ctx, *ct, P["w"], sl_all, rl_all, *args)
    assert data.shape == (nnt, ns_call * nx)
    one = [np.ascontiguousarray(data[:, k * nx:(k + 1) * nx]) for k in range(nshots)]
    vtiamd.objfunc_gradient2d(ctx, one[0], *ci, P["w"], sl[:1], rloc, *args)
    t0 = time.perf_counter()
    for k in range(nshots):
        vtiamd.objfunc_gradient2d(ctx, one[k], *ci, P["w"], sl[k:k + 1], rloc, *args)
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    vtiamd.objfunc_gradient2d(ctx, data, *ci, P["w"], sl_all, rl_all, *args)
    el32 = time.perf_counter() - t0
    return {"shots": nshots, "grid_padded": [nz + 2 * npad, nx + 2 * npad], "steps": nt - 1, "receivers": nx,
            "model": "cfg2 initial model (smoothed 1D gradient + 64 anomalies), data of the true model",
            "ms_per_shot": round(el * 1e3 / nshots, 2), "shot_gradients_per_s": round(nshots / el, 3),
            f"call_of_{ns_call}_shots": {"s": round(el32, 3), "shot_gradients_per_s": round(ns_call / el32, 3),
                                          "speedup_over_one_shot_calls": round((ns_call / el32) / (nshots / el), 2)}}


def fwi_lbfgs_2d(vtiamd, ctx, iters=5, nfg=20):
    """The cfg2 inversion as fwi-lbfgs-c-param.cu runs it (parameter vepsdel):
    32 shots of data of the true model, the initial model (v, eps, del scaled
    by v0 = 100 m/s, eps0 = del0 = 0.01, so that the first trial step, of
    length 1 along -g, moves v by at most 100 m/s), objFuncGradientVEpsDel on the GPU
    (one call with the 32 shots, run as one batch; wbottom 0) and the FP64
    L-BFGS (m = 5) with its vectors resident in HBM, until `iters`
    iterations are complete or `nfg` evaluations are spent.  The objective
    reaches L-BFGS scaled to 1e6 at the initial model, since the reference's
    stopping test |g|/max(1,|x|) <= 1e-5 is absolute.  Wall time of the
    whole loop; the objective of every evaluation."""
    from vtiamd import workloads as W

    ns = 32
    P = W.cfg2(nshots=ns)
    nx, nz, npad, nt = P["nx"], P["nz"], P["npad"], P["nt"]
    taper = vtiamd.init_taper2d(nx, nz, npad)
    geo = (nx, nz, nt, npad, P["dx"], P["dz"], P["dt"], P["rate"])
    sl_all = P["sloc"].copy()
    sl_all[:, 3] = np.arange(ns) * nx
    rl_all = np.tile(P["rloc"], (ns, 1))
    data = vtiamd.synthetic2d(ctx, *W.cij(*P["true"]), P["w"], sl_all, rl_all, taper, *geo)
    v0, eps0, del0 = 100.0, 0.01, 0.01
    v, e, d = P["init"]
    nn = v.size
    x = np.concatenate([(v / np.float32(v0)).ravel(), (e / np.float32(eps0)).ravel(),
                        (d / np.float32(del0)).ravel()]).astype(np.float32)
    opt = vtiamd.LBFGS(3 * nn, 5, ctx=ctx)
    fs, iflag, t_grad = [], 0, 0.0
    t0 = time.perf_counter()
    while True:
        t1 = time.perf_counter()
        f, g, e2, d2, _ = vtiamd.objfunc_gradient2d_vepsdel(ctx, data, x[:nn], x[nn:2 * nn], x[2 * nn:], v0, eps0,
                                                             del0, P["w"], sl_all, rl_all, taper, *geo)
        t_grad += time.perf_counter() - t1
        x[nn:2 * nn], x[2 * nn:] = e2, d2  # checkEpsDel projects eps/del in place, as in the reference
        fs.append(float(f))
        f0 = fs[0]  # the objective goes to L-BFGS scaled to 1e6 at the initial model (its stopping test is absolute)
        iflag = opt.step(x, f / f0 * 1e6, (np.asarray(g, np.float64) / f0 * 1e6).astype(np.float32).ravel(), iflag)
        if iflag != 1 or opt.iter > iters or len(fs) >= nfg:
            break
    wall = time.perf_counter() - t0
    done = min(opt.iter - 1, iters) if iflag == 1 else opt.iter
    opt.close()
    return {"shots": ns, "grid_padded": [nz + 2 * npad, nx + 2 * npad], "steps": nt - 1, "unknowns": 3 * nn,
            "lbfgs": "FP64, m=5, vectors in HBM (vti_lbfgs_create_device)", "iterations_completed": int(done),
            "evaluations": len(fs), "objective_initial": float(f"{fs[0]:.6e}"),
            "objective_relative": [round(q / fs[0], 6) for q in fs], "wall_s": round(wall, 2),
            "gradient_s": round(t_grad, 2), "shot_gradients_per_s": round(ns * len(fs) / t_grad, 2)}


def synthetic_cfg1(vtiamd, ctx, ns=16):
    """synthetic_f on the 256x256 (+40 boundary) grid, 2000 steps, 256
    receivers, rate 4 ms: 16 shots in one call (run as one batch) against
    one shot per call: shots/s of both."""
    nx, nz, npad, nt, dx, dt, rate = 256, 256, 40, 2001, 10.0, 1e-3, 4e-3
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    v = 1500.0 + 2500.0 * np.arange(nnz, dtype=np.float32)[:, None] / nnz * np.ones((1, nnx), np.float32)
    c33 = (v * v).astype(np.float32)
    c11, c13 = (c33 * 1.2).astype(np.float32), (c33 * 1.05).astype(np.float32)
    taper = vtiamd.init_taper2d(nx, nz, npad)
    t = np.arange(nt) * dt - 0.1
    a = (np.pi * 15.0 * t) ** 2
    w = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    rloc = np.array([[npad + i, npad + 1] for i in range(nx)], np.int32)
    sl = np.array([[npad + 8 + (k * (nx - 16)) // (ns - 1), npad + 2, nx, 0] for k in range(ns)], np.int32)
    args = (taper, nx, nz, nt, npad, dx, dx, dt, rate)
    vtiamd.synthetic2d(ctx, c11, c13, c33, w, sl[:2], rloc, *args)
    t0 = time.perf_counter()
    for k in range(4):
        vtiamd.synthetic2d(ctx, c11, c13, c33, w, sl[k:k + 1], rloc, *args)
    one = 4 / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    vtiamd.synthetic2d(ctx, c11, c13, c33, w, sl, rloc, *args)
    many = ns / (time.perf_counter() - t0)
    return {"grid_padded": [nnz, nnx], "steps": nt - 1, "receivers": nx, "one_shot_per_call_shots_per_s": round(one, 2),
            f"call_of_{ns}_shots_shots_per_s": round(many, 2), "speedup": round(many / one, 2)}


def cfg3_jobs(vtiamd, ctx, c11, c13, c33, npad, dx, dt, nt=3010):
    """The whole 3D jobs of the 460x572x572 configuration, one shot: forward
    modeling (modelDataCij3d_f) recording a 512 x 512 receiver grid two
    points below the boundary layer, and RTM (rtmCij3d_f): random-boundary
    forward, then the fused backward pass with the imaging condition.  nt =
    3010 makes the reference pipeline advance 3009 levels ((nt-2) % 64 == 0).
    The observed data of the RTM are the modeled traces."""
    nz, ny, nx = c11.shape
    rate = 4e-3
    w = ricker_np(nt, dt, 12.0, 0.1)
    i0 = (nx - 512) // 2
    zr = (npad + 2) * dx + 0.1
    ix, iy = np.meshgrid(np.arange(512) + i0, np.arange(512) + i0, indexing="xy")
    recloc = np.stack([ix.ravel() * dx + 0.1, iy.ravel() * dx + 0.1, np.full(ix.size, zr), np.ones(ix.size)],
                      axis=1).astype(np.float32)
    souloc = np.array([[nx // 2 * dx + 0.1, ny // 2 * dx + 0.1, (npad + 4) * dx + 0.1, ix.size, 0]], np.float32)
    o, d = (0.0, 0.0, 0.0), (dx, dx, dx)
    lmax = ((nt - 2) // 64) * 64 + 1
    npts = nz * ny * nx
    t0 = time.perf_counter()
    obs = vtiamd.model_data3d(ctx, souloc, np.arange(1), recloc, w, c11, c13, c33, nt, npad, o, d, dt, rate)
    tm = time.perf_counter() - t0
    rec3 = np.ascontiguousarray(recloc[:, :3])
    t0 = time.perf_counter()
    img = vtiamd.rtm3d(ctx, souloc, np.arange(1), rec3, w, c11, c13, c33, nt, npad, o, d, dt, rate, obs)
    tr = time.perf_counter() - t0
    assert np.isfinite(img).all() and np.abs(img).max() > 0
    # algorithmic bytes: forward 36 B/pt-level; RTM 36 (random-boundary forward) + 68 (fused backward)
    return {"grid_padded": [nz, ny, nx], "levels": lmax, "receivers": int(ix.size), "samples": int(obs.size // ix.size),
            "modeling_s": round(tm, 3), "modeling_gpts_s": round(npts * lmax / tm / 1e9, 2),
            "rtm_s": round(tr, 3), "rtm_floor_s_at_8tbs": round((36 + 68) * npts * lmax / 8e12, 3),
            "rtm_level_gpts_s": round(npts * lmax / tr / 1e9, 2),
            "note": "host-side times of the whole calls: uploads, the 790 MB trace transfer and setup included"}


def ricker_np(nt, dt, f, t0):
    t = np.arange(nt) * dt - t0
    a = (np.pi * f * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def run_sections(extra, sections):
    """Run the secondary workloads in order, each result under its name.
    An error is recorded in place of that section's result (its traceback
    goes to stderr) and the later sections are marked as not run, since the
    device may be unusable after it; the headline measured before them is
    still reported.  Returns the name of the failed section, or None."""
    failed = None
    for name, fn in sections:
        if failed is not None:
            extra[name] = {"not_run": f"after the error in {failed}"}
            continue
        progress(f"section {name}")
        try:
            extra[name] = fn()
            progress(f"section {name} done")
        except Exception as e:  # reported in the output line
            traceback.print_exc(file=sys.stderr)
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
            failed = name
    return failed


def serve_trace(vtiamd, nreq=40, rate=5.0, frac3d=0.3, seed=5, surveys=False):
    """A seeded request trace: Poisson arrivals at `rate` req/s, a fraction
    frac3d of 3D modeling requests (each dimension 128-400, 1000-3000
    steps), the rest 2D gradient requests (nz 256-512, nx 256-1024 + a
    40-point boundary, 1000-3000 steps), on seeded layered models (four 2D
    and three 3D grids) that the server keeps resident.  surveys: the 2D
    requests are single shots of two surveys instead -- the first two 2D
    grids, each with its own fixed step count (2001, 1501), a source at a
    random column -- as an inversion service sees them."""
    rng = np.random.default_rng(seed)
    npad2 = 40
    grids2 = sorted({(int(rng.integers(256, 513)), int(rng.integers(256, 1025))) for _ in range(4)})
    grids3 = [tuple(int(v) for v in rng.integers(128, 401, 3)) for _ in range(3)]
    models2 = {}
    for nz, nx in grids2:
        nnz, nnx = nz + 2 * npad2, nx + 2 * npad2
        z = np.arange(nnz, dtype=np.float32)[:, None] * np.ones((1, nnx), np.float32)
        v = 1500.0 + 3000.0 * z / nnz + 80.0 * np.sin(np.arange(nnx, dtype=np.float32)[None, :] / 53.0)
        c33 = (v * v).astype(np.float32)
        models2[(nz, nx)] = (c33 * np.float32(1.2), c33 * np.float32(1.05), c33, vtiamd.init_taper2d(nx, nz, npad2))
    models3 = {g: layered(*g, seed=sum(g))[:3] for g in grids3} if frac3d > 0 else {}
    reqs, t = [], 0.0
    for _ in range(nreq):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() >= frac3d:
            if surveys:
                k = int(rng.integers(2))
                g = grids2[k]
                reqs.append({"t": t, "kind": "2d", "grid": g, "nt": (2001, 1501)[k],
                             "ix": npad2 + int(rng.integers(8, g[1] - 8))})
                continue
            g = grids2[int(rng.integers(len(grids2)))]
            reqs.append({"t": t, "kind": "2d", "grid": g, "nt": 4 * int(rng.integers(250, 751)) + 1})
        else:
            g = grids3[int(rng.integers(len(grids3)))]
            reqs.append({"t": t, "kind": "3d", "grid": g, "nt": int(rng.integers(1000, 3001))})
    return {"reqs": reqs, "grids2": grids2, "grids3": grids3, "models2": models2, "models3": models3,
            "npad2": npad2, "rate": rate, "frac3d": frac3d}


def serve_run(vtiamd, tr, slots, realtime, batch=8):
    """Run a trace through a server with `slots` workers merging up to
    `batch` 2D requests per run (1: none); realtime: submit at the arrival
    times, else all at once.  Returns latency/throughput stats."""
    npad2, dx2, dt, rate_s = tr["npad2"], 10.0, 1e-3, 4e-3

    def wavelet(nt):
        tt = np.arange(nt) * dt - 0.1
        a = (np.pi * 12.0 * tt) ** 2
        return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)

    def geom3(g):
        nz, ny, nx = g
        d = 12.5
        rec = [(ix * d + 0.1, iy * d + 0.1, 32 * d + 0.1, 1.0) for iy in (ny // 3, 2 * ny // 3)
               for ix in range(30, nx - 30, 4)]
        sou = np.array([[nx // 2 * d + 0.1, ny // 2 * d + 0.1, 34 * d + 0.1, len(rec), 0]], np.float32)
        return sou, np.array(rec, np.float32)

    with vtiamd.Server(0, slots=slots, batch=batch) as srv:
        h2 = {g: srv.put_model(*tr["models2"][g][:3]) for g in tr["grids2"]}
        h3 = {g: srv.put_model(*tr["models3"][g]) for g in tr["models3"]}
        tickets = []
        t0 = time.perf_counter()
        for r in tr["reqs"]:
            if realtime:
                time.sleep(max(0.0, t0 + r["t"] - time.perf_counter()))
            if r["kind"] == "2d":
                nz, nx = r["grid"]
                nnt = (r["nt"] - 1) // 4 + 1
                sl = np.array([[r.get("ix", npad2 + nx // 2), npad2 + 2, nx, 0]], np.int32)
                rl = np.array([[npad2 + i, npad2 + 1] for i in range(nx)], np.int32)
                tickets.append(srv.grad2d(h2[r["grid"]], np.zeros((nnt, nx), np.float32), wavelet(r["nt"]), sl, rl,
                                          tr["models2"][r["grid"]][3], nx, nz, r["nt"], npad2, dx2, dx2, dt, rate_s))
            else:
                sou, rec = geom3(r["grid"])
                tickets.append(srv.model3d(h3[r["grid"]], sou, np.arange(1), rec, wavelet(r["nt"]), r["grid"],
                                           r["nt"], 30, (0.0, 0.0, 0.0), (12.5, 12.5, 12.5), dt, rate_s))
        for tk in tickets:
            srv.wait(tk)
        st = srv.stats()
    return {"p50_s": round(st["p50_s"], 4), "p99_s": round(st["p99_s"], 4), "req_per_s": round(st["req_per_s"], 3),
            "peak_gb": round(st["peak_bytes"] / 2**30, 2), "merged_requests": st["merged_requests"]}


def serve_stream(vtiamd, nreq=300, nsat=150):
    """The served stream of the cfg5 mix (70% 2D gradients, 30% 3D modeling)
    at its Poisson rate of 5 req/s, shortened from 10 minutes to `nreq`
    requests (about a minute) so that p99 is the 3rd-slowest of 300 rather
    than the maximum, through a 4-slot server and through one slot without
    merging (each request alone, in arrival order: the plain per-request
    launch); the saturation throughput (all requests at once) of both on
    `nsat` of them, also for a stream of 2D gradient requests only; then the
    same for a stream whose 2D requests are single shots of two surveys,
    which the server merges into batched runs (12 req/s, `nreq` requests)."""
    out = {"requests": nreq, "mix": "70% 2D gradient, 30% 3D modeling",
           "duration_note": f"{nreq} Poisson arrivals at 5 req/s (about {nreq / 5:.0f} s) in place of 10 minutes"}
    tr = serve_trace(vtiamd, nreq, 5.0)
    out["grids_2d"] = [list(g) for g in tr["grids2"]]
    out["grids_3d"] = [list(g) for g in tr["grids3"]]
    served, plain = serve_run(vtiamd, tr, 4, True), serve_run(vtiamd, tr, 1, True, batch=1)
    out["poisson_5rps"] = {"served_4slots": served, "plain_1slot": plain,
                           "p99_ratio_plain_over_served": round(plain["p99_s"] / max(served["p99_s"], 1e-9), 2)}
    progress("served stream: cfg5 mix at 5 req/s")

    def saturation(tr):
        sat = {"served_4slots": serve_run(vtiamd, tr, 4, False)["req_per_s"],
               "plain_1slot": serve_run(vtiamd, tr, 1, False, batch=1)["req_per_s"]}
        sat["ratio"] = round(sat["served_4slots"] / max(sat["plain_1slot"], 1e-9), 2)
        return sat

    out["saturation_req_s"] = saturation(serve_trace(vtiamd, nsat, 5.0))
    out["saturation_requests"] = nsat
    out["saturation_2d_only_shot_gradients_per_s"] = saturation(serve_trace(vtiamd, nsat, 5.0, frac3d=0.0, seed=6))
    progress("served stream: saturation")
    sv = {"mix": "70% single-shot 2D gradients of two surveys (grids_2d[0] at 2000 steps, grids_2d[1] at 1500), "
                 "30% 3D modeling", "requests": nreq}
    tr = serve_trace(vtiamd, nreq, 12.0, surveys=True, seed=7)
    served, plain = serve_run(vtiamd, tr, 4, True), serve_run(vtiamd, tr, 1, True, batch=1)
    sv["poisson_12rps"] = {"served_4slots": served, "plain_1slot": plain,
                           "p99_ratio_plain_over_served": round(plain["p99_s"] / max(served["p99_s"], 1e-9), 2)}
    sv["saturation_req_s"] = saturation(serve_trace(vtiamd, nsat, 12.0, surveys=True, seed=9))
    tr = serve_trace(vtiamd, nsat, 5.0, frac3d=0.0, seed=8, surveys=True)
    sat = saturation(tr)
    sat["served_2slots"] = serve_run(vtiamd, tr, 2, False)["req_per_s"]
    sat["ratio_2slots"] = round(sat["served_2slots"] / max(sat["plain_1slot"], 1e-9), 2)
    sv["saturation_2d_only_shot_gradients_per_s"] = sat
    sv["saturation_2d_only_requests"] = nsat
    out["surveys"] = sv
    progress("served stream: surveys")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--small", action="store_true", help="small grid (profiling runs)")
    ap.add_argument("--no-extra", action="store_true", help="only the headline kernel (profiling runs)")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D and serving workloads (profiling runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph replays of the batched workload (kernel-trace profiling runs)")
    ap.add_argument("--only", choices=["shot2d", "batch2d", "serve", "jobs3d", "fwi2d"],
                    help="run one secondary workload alone and print its JSON (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: this path shards by shot across GPUs (one RCCL all-reduce of the "
            "gradient per iteration) and nothing here runs on more than one device.",
            file=sys.stderr,
        )
        sys.exit(2)

    import vtiamd

    if args.only:
        with vtiamd.Context(0) as ctx:
            if args.only == "shot2d":
                r = shot_gradient_2d(vtiamd, ctx)
                r["synthetic_256x256_2000steps"] = synthetic_cfg1(vtiamd, ctx)
            elif args.only == "batch2d":
                r = batch_gradient_2d(vtiamd, ctx, graph=not args.no_graph)
            elif args.only == "fwi2d":
                r = fwi_lbfgs_2d(vtiamd, ctx)
            elif args.only == "jobs3d":
                c11, c13, c33, _ = layered(460, 572, 572)
                r = cfg3_jobs(vtiamd, ctx, c11, c13, c33, 30, 12.5, 1e-3)
            else:
                r = serve_stream(vtiamd)
        print(json.dumps({"workload": args.only, **r}), flush=True)
        return

    nz, ny, nx, npad = (460, 572, 572, 30) if not args.small else (132, 140, 140, 10)
    dx, dt = 12.5, 1.0e-3
    c11, c13, c33, ved = layered(nz, ny, nx)
    npts = nz * ny * nx
    workload = f"3d_vti_forward_step_{nz}x{ny}x{nx}"

    ctx = vtiamd.Context(0)
    prop = vtiamd.Prop3D(ctx, nx, ny, nz, npad, dx, dx, dx, dt)
    if args.zchunk:
        prop.set_zchunk(args.zchunk)
    prop.set_model(c11, c13, c33)
    sou = nx // 2 + (ny // 2) * nx + (npad + 10) * nx * ny
    prop.reset(sou, 1.0e-6)
    prop.forward_steps(args.warmup, absorbing=True)
    ctx.sync()
    t0 = time.perf_counter()
    kernel_ms = prop.forward_steps(args.steps, absorbing=True, timed=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / max(args.steps, 1)
    progress(f"headline: {args.steps} forward steps, {ms_per_step:.4f} ms/step")
    value = npts / (ms_per_step * 1e-3) / 1e9
    achieved = BYTES_FWD * npts / (kernel_ms * 1e-3) / 1e9
    traffic = load_traffic("fwd3d_kernel", workload)
    out = {
        "metric": "3D VTI step throughput (Gpts/s) and HBM-roofline fraction",
        "value": round(value, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered dipping VTI model, no checkpoint)",
        "config": {
            "workload": workload,
            "grid_padded": [nz, ny, nx],
            "npad": npad,
            "dx_m": dx,
            "dt_s": dt,
            "absorbing": True,
            "device_bytes": prop.nbytes,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_ms": round(kernel_ms, 4),
            "bytes_per_point": BYTES_FWD,
        },
    }
    # secondary kernels of the same path on the same grid (reported, not the headline)
    extra = {}
    if args.no_extra:
        prop.close()
        ctx.close()
        print(json.dumps(out), flush=True)
        return
    bw = prop.backward_steps(max(1, args.steps // 2), mode=0, timed=True)
    extra["bwd_gradient_step"] = {
        "ms": round(bw, 4), "gpts_s": round(npts / (bw * 1e-3) / 1e9, 3), "bytes_per_point": BYTES_BWD,
        "achieved_gbs": round(BYTES_BWD * npts / (bw * 1e-3) / 1e9, 1),
    }
    br = prop.backward_steps(max(1, args.steps // 2), mode=1, timed=True)
    extra["bwd_rtm_step"] = {
        "ms": round(br, 4), "gpts_s": round(npts / (br * 1e-3) / 1e9, 3), "bytes_per_point": BYTES_RTM,
        "achieved_gbs": round(BYTES_RTM * npts / (br * 1e-3) / 1e9, 1),
    }
    progress("backward steps timed")
    prop.set_arith(0)
    f32ms = prop.forward_steps(max(1, args.steps // 2), absorbing=True, timed=True)
    extra["fwd_step_fp32_arith"] = {"ms": round(f32ms, 4), "gpts_s": round(npts / (f32ms * 1e-3) / 1e9, 3)}
    prop.close()
    failed = None
    if not args.small and not args.no_2d:
        failed = run_sections(extra, [
            ("3d_jobs_460x572x572_3009_levels", lambda: cfg3_jobs(vtiamd, ctx, c11, c13, c33, npad, dx, dt)),
            ("2d_shot_gradient_512x1024_3000steps", lambda: shot_gradient_2d(vtiamd, ctx)),
            ("2d_fwi_lbfgs_512x1024_32shots", lambda: fwi_lbfgs_2d(vtiamd, ctx)),
            ("2d_synthetic_256x256_2000steps", lambda: synthetic_cfg1(vtiamd, ctx)),
            ("2d_batched_vs_plain_gradient", lambda: batch_gradient_2d(vtiamd, ctx)),
            ("served_stream", lambda: serve_stream(vtiamd)),
        ])
    out["extra"] = extra
    if failed is not None:
        out["errors"] = [f"{failed}: {extra[failed]['error']}"]
    if not args.no_cpu and not args.small:
        progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(ved, npad, dx, dt)
    try:
        ctx.close()
    except Exception:
        if failed is None:
            raise
        traceback.print_exc(file=sys.stderr)
    print(json.dumps(out), flush=True)
    if failed is not None:
        sys.exit(3)  # the headline is in the line; a section failed


if __name__ == "__main__":
    main()
