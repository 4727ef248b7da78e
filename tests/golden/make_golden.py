#!/usr/bin/env python3
"""Writes tests/golden/reference_host.json from the compiled reference pieces.

Run in a container that holds the reference tree and after `make -C oracle ref`.
It executes, as built from the reference's own sources under oracle/_ref/:
  - randomBound.x  (lib/random-boundary/randomBound.cpp + lib/io/myio.cpp):
    the reference's random-boundary program, run in its own process on SEP
    files this script writes;
  - libref.so: VEpsDel2Cij / dVEpsDel2dCij / GradCij2GradVEpsDel
    (lib/vti-fwi-2d/conversions.cpp), checkEpsDel (lib/vti-fwi-2d/check.cpp),
    init_abc 2D (lib/vti-fwi-2d/init.cpp),
    getBoundary / putBoundary / zeroBoundary (lib/vti-fwi-3d/gpu/boundary.cpp),
    ricker (lib/vti-fwi-2d/ricker.cpp);
  - libreflbfgs.so: lbfgs (lib/lbfgs/lbfgs_c.cpp, lsearch.cpp, la.cpp) on the
    test functions of lib/lbfgs/functions.cpp, driven as lib/lbfgs/test1c.cpp
    drives them.
Inputs are generated here from seeds; the fixture stores the seeds, shapes,
SHA-256 of every output array and a few sampled values.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.join(HERE, "reference_host.json")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def samples(a: np.ndarray, k=8):
    flat = np.ascontiguousarray(a).ravel()
    idx = np.linspace(0, flat.size - 1, k).astype(int)
    return {"index": idx.tolist(), "value": [float(flat[i]) for i in idx]}


def rb_input(dims, seed):
    """Deterministic input velocity for the random-boundary fixtures."""
    n1, n2, n3 = dims
    rng = np.random.default_rng(seed)
    z = np.arange(n3 * n2 * n1, dtype=np.float64).reshape(n3, n2, n1)
    base = 1500.0 + 2500.0 * rng.random() + 0.01 * (z % 997)
    return base.astype(np.float32)


def run_random_boundary(dims, pad, pctG, seed, input_seed):
    n1, n2, n3 = dims
    vel = rb_input(dims, input_seed)
    with tempfile.TemporaryDirectory() as d:
        binf = os.path.join(d, "velin.bin")
        vel.tofile(binf)
        hin = os.path.join(d, "velin.H")
        with open(hin, "w") as f:
            f.write(f"n1={n1} o1=0 d1=1\nn2={n2} o2=0 d2=1\nn3={n3} o3=0 d3=1\n")
            f.write(f"in={binf}\ndata_format=native_float\n")
        hout = os.path.join(d, "velout.H")
        subprocess.run(
            [os.path.join(REF, "randomBound.x"), f"modelIn={hin}", f"modelOut={hout}", f"seed={seed}",
             f"pad={pad}", f"pctG={pctG}", f"datapath={d}/"],
            check=True, capture_output=True,
        )
        out = np.fromfile(os.path.join(d, "velout.H@"), dtype=np.float32)
    assert out.size == vel.size
    return vel, out.reshape(vel.shape)


def main():
    if not os.path.exists(os.path.join(REF, "randomBound.x")):
        sys.exit("oracle/_ref not built: make -C oracle ref")
    fx = {"generator": "tests/golden/make_golden.py", "random_boundary": [], "conversions": {}, "misc": {}}

    cases = [
        ((336, 336, 1), 40, 0.99, 2017, 1),  # 2D, the 256x256 + 40-point configuration
        ((336, 336, 1), 40, 0.9, 7, 2),
        ((70, 64, 60), 10, 0.99, 2017, 3),  # 3D
        ((72, 80, 56), 12, 0.95, 123, 4),
    ]
    for dims, pad, pctG, seed, iseed in cases:
        vin, vout = run_random_boundary(dims, pad, pctG, seed, iseed)
        fx["random_boundary"].append({
            "dims": list(dims), "pad": pad, "pctG": pctG, "seed": seed, "input_seed": iseed,
            "sha256_in": sha(vin), "sha256_out": sha(vout), "samples": samples(vout),
            "changed_points": int(np.count_nonzero(vin != vout)),
        })
        print("random boundary", dims, pad, pctG, seed, "changed", fx["random_boundary"][-1]["changed_points"])

    L = ctypes.CDLL(os.path.join(REF, "libref.so"))
    P = ctypes.c_void_p
    f = ctypes.c_float
    n = 4096
    rng = np.random.default_rng(11)
    v = rng.uniform(0.4, 1.5, n).astype(np.float32)
    eps = rng.uniform(0.0, 0.3, n).astype(np.float32)
    dl = rng.uniform(-0.05, 0.2, n).astype(np.float32)
    dv = rng.standard_normal(n).astype(np.float32)
    de = rng.standard_normal(n).astype(np.float32)
    dd = rng.standard_normal(n).astype(np.float32)
    g = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    v0, e0, d0 = 2500.0, 0.8, 1.3
    c = [np.zeros(n, np.float32) for _ in range(6)]
    L.ref_VEpsDel2Cij.argtypes = [P] * 6 + [f, f, f, ctypes.c_size_t]
    L.ref_VEpsDel2Cij(*[a.ctypes.data for a in c[:3]], v.ctypes.data, eps.ctypes.data, dl.ctypes.data, v0, e0, d0, n)
    fx["conversions"]["VEpsDel2Cij"] = [sha(a) for a in c[:3]]
    L.ref_dVEpsDel2dCij.argtypes = [P] * 12 + [f, f, f, ctypes.c_size_t]
    L.ref_dVEpsDel2dCij(*[a.ctypes.data for a in c], v.ctypes.data, eps.ctypes.data, dl.ctypes.data,
                        dv.ctypes.data, de.ctypes.data, dd.ctypes.data, v0, e0, d0, n)
    fx["conversions"]["dVEpsDel2dCij"] = [sha(a) for a in c]
    gout = [np.zeros(n, np.float32) for _ in range(3)]
    L.ref_GradCij2GradVEpsDel.argtypes = [P] * 9 + [f, f, f, ctypes.c_size_t]
    L.ref_GradCij2GradVEpsDel(*[a.ctypes.data for a in gout], *[a.ctypes.data for a in g], v.ctypes.data,
                              eps.ctypes.data, dl.ctypes.data, v0, e0, d0, n)
    fx["conversions"]["GradCij2GradVEpsDel"] = [sha(a) for a in gout]
    fx["conversions"]["inputs"] = {"seed": 11, "n": n, "v0": v0, "eps0": e0, "del0": d0}

    # checkEpsDel (lib/vti-fwi-2d/check.cpp): eps/del pairs with del0*del above, equal to and below eps0*eps
    rng2 = np.random.default_rng(12)
    ce = rng2.uniform(0.0, 0.3, n).astype(np.float32)
    cd = rng2.uniform(0.0, 0.3, n).astype(np.float32)
    cd[::7] = ce[::7]
    cm = np.zeros(n, np.float32)
    L.ref_checkEpsDel.argtypes = [P, P, f, f, ctypes.c_size_t, P]
    L.ref_checkEpsDel(ce.ctypes.data, cd.ctypes.data, 1.0, 1.0, n, cm.ctypes.data)
    fx["conversions"]["checkEpsDel"] = {"seed": 12, "n": n, "eps0": 1.0, "del0": 1.0, "equal_every": 7,
                                        "sha256": [sha(ce), sha(cd), sha(cm)],
                                        "projected": int(np.count_nonzero(cm == 2))}

    for nx, nz, npad in ((256, 256, 40), (1024, 512, 40), (30, 20, 7)):
        t = np.zeros((nz + 2 * npad) * (nx + 2 * npad), np.float32)
        L.ref_init_abc2d.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ref_init_abc2d(t.ctypes.data, nx, nz, npad)
        fx["misc"][f"init_abc2d_{nx}_{nz}_{npad}"] = sha(t)

    nx, ny, nz, npad = 23, 19, 17, 4
    vol = rng.standard_normal(nx * ny * nz).astype(np.float32)
    nb = nx * ny * nz - (nx - 2 * npad) * (ny - 2 * npad) * (nz - 2 * npad)
    b = np.zeros(nb, np.float32)
    L.ref_getBoundary.argtypes = [P, P] + [ctypes.c_int] * 4
    L.ref_getBoundary(b.ctypes.data, vol.ctypes.data, nx, ny, nz, npad)
    fx["misc"]["getBoundary"] = {"dims": [nx, ny, nz, npad], "seed": 11, "sha256": sha(b)}
    vol2 = vol.copy()
    L.ref_zeroBoundary.argtypes = [P] + [ctypes.c_int] * 4
    L.ref_zeroBoundary(vol2.ctypes.data, nx, ny, nz, npad)
    fx["misc"]["zeroBoundary"] = sha(vol2)

    for freq, nt, dt, ot, tdelay, s in ((15.0, 2000, 1e-3, 0.0, 0.1, 1.0), (10.0, 3001, 1e-3, 0.0, 0.12, 2.5),
                                        (12.0, 3000, 1e-3, -0.05, 0.1, 1.0)):
        w = np.zeros(nt, np.float32)
        L.ref_ricker.argtypes = [P, f, ctypes.c_int, f, f, f, f]
        L.ref_ricker(w.ctypes.data, freq, nt, dt, ot, tdelay, s)
        fx["misc"][f"ricker_{freq}_{nt}_{dt}_{ot}_{tdelay}_{s}"] = sha(w)

    # the reference L-BFGS (lib/lbfgs, built without OpenMP) on its own test
    # functions in the loop of lib/lbfgs/test1c.cpp (at most 201 calls)
    LB = ctypes.CDLL(os.path.join(REF, "libreflbfgs.so"))
    LB.ref_lbfgs_run.argtypes = [ctypes.c_int, ctypes.c_int, P, P, ctypes.POINTER(ctypes.c_int)]
    fx["lbfgs"] = {}
    for func, n in ((0, 2), (1, 2), (2, 100), (3, 100), (4, 50), (5, 25)):
        xo = np.zeros(n, np.float32)
        ft = np.zeros(201, np.float32)
        fl = ctypes.c_int()
        saved = os.dup(2)
        os.dup2(os.open(os.devnull, os.O_WRONLY), 2)  # the reference reports every step on stderr
        try:
            calls = LB.ref_lbfgs_run(func, 201, xo.ctypes.data, ft.ctypes.data, ctypes.byref(fl))
        finally:
            os.dup2(saved, 2)
        fx["lbfgs"][str(func)] = {"n": n, "calls": calls, "iflag": fl.value, "x": [float(v) for v in xo],
                                  "f": [float(v) for v in ft[:calls]]}
        print("lbfgs", func, "calls", calls, "iflag", fl.value)

    # a small dataset written by the reference's own lib/io (through
    # randomBound.x): header + binary, read back by vtiamd.sep in the tests
    sepdir = os.path.join(HERE, "sep")
    os.makedirs(sepdir, exist_ok=True)
    for f in os.listdir(sepdir):
        os.remove(os.path.join(sepdir, f))
    vin = rb_input((48, 40, 1), 5)
    vin.tofile(os.path.join(sepdir, "velin.bin"))
    hin = os.path.join(sepdir, "velin.H")
    with open(hin, "w") as f:
        f.write("n1=48 o1=0 d1=10\nn2=40 o2=5 d2=12.5\nn3=1 o3=0 d3=1\n")
        f.write(f"in={os.path.join(sepdir, 'velin.bin')}\ndata_format=native_float\n")
    subprocess.run([os.path.join(REF, "randomBound.x"), f"modelIn={hin}", f"modelOut={sepdir}/velout.H", "seed=9",
                    "pad=8", "pctG=0.9", f"datapath={sepdir}/"], check=True, capture_output=True, timeout=120)
    fx["sep"] = {"header": "sep/velout.H", "dims": [48, 40, 1], "pad": 8, "pctG": 0.9, "seed": 9, "input_seed": 5,
                 "sha256": sha(np.fromfile(os.path.join(sepdir, "velout.H@"), np.float32))}
    print("sep dataset", fx["sep"]["sha256"][:12])

    with open(OUT, "w") as fo:
        json.dump(fx, fo, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
