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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "thesis-source-code_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_FWD = 36  # sigma_x, sigma_z cur+prev, c11, c13, c33 read; sigma_x, sigma_z written
BYTES_BWD = 84  # fused gradient step: 14 reads + 7 writes of 4 bytes per point
BYTES_RTM = 68  # fused RTM step: 12 reads + 5 writes


def layered(nz, ny, nx, seed=2024):
    rng = np.random.default_rng(seed)
    nl = 6
    z = np.arange(nz, dtype=np.float32)[:, None, None]
    y = np.arange(ny, dtype=np.float32)[None, :, None]
    x = np.arange(nx, dtype=np.float32)[None, None, :]
    depth = z + 0.12 * (x - nx / 2) + 0.06 * (y - ny / 2)
    bounds = np.sort(rng.uniform(0.1, 0.9, nl - 1)) * nz
    layer = np.searchsorted(bounds, depth).astype(np.int8)
    vl = np.sort(rng.uniform(1500.0, 4500.0, nl)).astype(np.float32)
    el = rng.uniform(0.0, 0.25, nl).astype(np.float32)
    dl = np.minimum(rng.uniform(0.0, 0.12, nl), el).astype(np.float32)
    v, e, d = vl[layer], el[layer], dl[layer]
    c33 = v * v
    c11 = c33 * (1 + 2 * e)
    c13 = c33 * np.sqrt(1 + 2 * d)
    return c11.astype(np.float32), c13.astype(np.float32), c33.astype(np.float32)


def load_traffic(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(f"{kernel}@{workload}")
    except (OSError, ValueError):
        return None


def cpu_baseline(c11, c13, c33, npad, dx, dt, budget_s=15.0):
    """Time the oracle's C/OpenMP step on the same grid for a bounded number of steps."""
    import oracle

    nz, ny, nx = c11.shape
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    rng = np.random.default_rng(0)
    st = [np.zeros((nz, ny, nx), np.float32) for _ in range(4)]
    st[0][:] = rng.standard_normal((nz, ny, nx), dtype=np.float32) * 1e-3
    t0 = time.perf_counter()
    oracle.steps3d(oracle.REF, st, c11, c13, c33, npad, dx, dx, dx, dt, 1, absorbing=True)
    t1 = time.perf_counter()
    one = t1 - t0
    n = max(1, min(30, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    oracle.steps3d(oracle.REF, st, c11, c13, c33, npad, dx, dx, dx, dt, n, absorbing=True)
    el = time.perf_counter() - t0
    gpts = nz * ny * nx * n / el / 1e9
    return {
        "value": round(gpts, 4),
        "unit": "Gpts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} forward steps (absorbing) of the full {nz}x{ny}x{nx} grid, "
        f"oracle C/OpenMP restatement in the reference arithmetic, {el:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--small", action="store_true", help="small grid (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: this path shards by shot across GPUs (one RCCL all-reduce of the "
            "gradient per iteration) and nothing here runs on more than one device.",
            file=sys.stderr,
        )
        sys.exit(2)

    import vtiamd

    nz, ny, nx, npad = (460, 572, 572, 30) if not args.small else (132, 140, 140, 10)
    dx, dt = 12.5, 1.0e-3
    c11, c13, c33 = layered(nz, ny, nx)
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
    prop.set_arith(0)
    f32ms = prop.forward_steps(max(1, args.steps // 2), absorbing=True, timed=True)
    extra["fwd_step_fp32_arith"] = {"ms": round(f32ms, 4), "gpts_s": round(npts / (f32ms * 1e-3) / 1e9, 3)}
    out["extra"] = extra
    prop.close()
    if not args.no_cpu and not args.small:
        out["cpu_baseline"] = cpu_baseline(c11, c13, c33, npad, dx, dt)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
