"""Independent checks of the oracle, as functions returning the measured error
(test infrastructure).  test_oracle_pins_history.py asserts that each error is
at roundoff; test_oracle_mutants.py runs the same functions against broken
builds of the oracle and asserts that each build fails at least one of them.
"""
import numpy as np

import brute_history as bh
from common3d import cij_from, layered_model, rel, ricker


# ---------------------------------------------------------------------------
# 3D: a 24 x 20 x 22 grid, nt = 66 (the pipeline then advances nt-1 levels,
# so the random-boundary reversal pairs levels exactly as the stored history).
# dt = 2^-10 s and 16 m spacing make dt2/d2 exact in float: the reference's
# forward step scales by float(dt2/d2) and its reversal by dt2 and 1/d2
# separately, which agree only then (otherwise the reversal drifts at the
# float rounding of dt2/d2, ~1e-7 here).
# ---------------------------------------------------------------------------
def problem3(nt=66, seed=3):
    nz, ny, nx, npad = 24, 20, 22, 5
    d = (16.0, 16.0, 16.0)
    o = (0.0, 0.0, 0.0)
    dt = 2.0 ** -10
    rate = 4 * dt
    v, e, dl = layered_model(nz, ny, nx, seed=seed, vmin=1800.0, vmax=3200.0)
    cij = np.stack(cij_from(v, e, dl)).astype(np.float32)
    vt = v.copy()
    vt[10:16, 7:13, 8:15] *= 1.05
    ctrue = np.stack(cij_from(vt, e, dl)).astype(np.float32)
    w = ricker(nt, dt, 25.0, 0.02)
    recs = []
    for iy in (10, 12):
        for ix in range(5, 17):
            recs.append((ix * d[0] + 0.1, iy * d[1] + 0.1, 9 * d[2] + 0.1, 1.0))
    recloc = np.array(recs, np.float32)
    recloc[3, 3] = 0.0  # a dead trace
    nr = len(recs)
    souloc = np.array([[11 * d[0] + 0.1, 10 * d[1] + 0.1, 14 * d[2] + 0.1, nr, 0]], np.float32)
    nbnd = nx * ny * nz - (nx - 2 * npad) * (ny - 2 * npad) * (nz - 2 * npad)
    rng = np.random.default_rng(seed + 1)
    shell = rng.uniform(0.75, 1.0, nbnd).astype(np.float32)
    return dict(nx=nx, ny=ny, nz=nz, npad=npad, d=d, o=o, dt=dt, rate=rate, nt=nt, cij=cij, ctrue=ctrue, w=w,
                recloc=recloc, souloc=souloc, nbnd=nbnd, shell=shell)


def _shells(orc, P):
    """Absorbing shell (the model's own) and random shell (scaled), packed in
    getBoundary order, and the model with the random shell put back."""
    nb = P["nbnd"]
    padb = np.zeros(3 * nb, np.float32)
    for k in range(3):
        orc.lib().orc3d_boundary(0, padb[k * nb:].ctypes.data, P["cij"][k].ctypes.data, P["nx"], P["ny"], P["nz"],
                                 P["npad"])
    randb = (padb * np.tile(P["shell"], 3)).astype(np.float32)
    crnd = P["cij"].copy()
    for k in range(3):
        orc.lib().orc3d_boundary(1, randb[k * nb:].ctypes.data, crnd[k].ctypes.data, P["nx"], P["ny"], P["nz"],
                                 P["npad"])
    return padb, randb, crnd


def _obs3(P, damping):
    """Observed data: traces of the true model (brute-force forward, absorbing)."""
    g = bh.Geo3(P["nz"], P["ny"], P["nx"], P["d"], P["dt"])
    sts = int(round(P["rate"] / P["dt"]))
    nnt = (P["nt"] - 1) // sts + 1
    geo = bh.geometry3(P["souloc"], 0, P["recloc"], 4, P["nx"], P["ny"], P["o"], P["d"])
    tr, _ = bh.forward3(g, [a.astype(np.float64) for a in P["ctrue"]], P["w"], geo, bh.lmax3(P["nt"]),
                        damping=damping, sts=sts, nnt=nnt)
    return tr.astype(np.float32).ravel()


def gradient3d_history(orc, tpow=0.0):
    """Oracle FP64 objFuncGradientCij3d_f against the stored-history gradient.
    Returns (relative L2 error of the gradient, relative error of J)."""
    P = problem3()
    damping = orc.init_damping3d(P["nx"], P["ny"], P["nz"], P["npad"])
    padb, randb, crnd = _shells(orc, P)
    obs = _obs3(P, damping)
    f, g = orc.gradient3d(orc.F64, P["souloc"], np.array([0]), P["recloc"], P["w"], P["cij"].copy(), padb, randb,
                          P["nt"], P["npad"], P["o"], P["d"], P["dt"], P["rate"], obs, tpow=tpow)
    J, gb = bh.gradient3(P["cij"], None, crnd, P["w"], P["souloc"], [0], P["recloc"], P["o"], P["d"], P["dt"],
                         P["nt"], P["rate"], obs, damping, tpow=tpow)
    return rel(g, gb), abs(f - J) / abs(J)


def rtm3d_history(orc):
    """Oracle FP64 rtmCij3d_f against the stored-history image."""
    P = problem3()
    damping = orc.init_damping3d(P["nx"], P["ny"], P["nz"], P["npad"])
    _, _, crnd = _shells(orc, P)
    obs = _obs3(P, damping)
    rec3 = np.ascontiguousarray(P["recloc"][:, :3])
    img = orc.rtm3d(orc.F64, P["souloc"], np.array([0]), rec3, P["w"], *crnd, P["nt"], P["npad"], P["o"], P["d"],
                    P["dt"], P["rate"], obs)
    ib = bh.rtm3(crnd, P["w"], P["souloc"], [0], rec3, P["o"], P["d"], P["dt"], P["nt"], P["rate"], obs, damping)
    return rel(img, ib)


def taylor3d(orc, k, hs=(0.5, 0.25)):
    """Central differences of the oracle's FP64 3D objective along a smooth
    blob in coefficient k (0: c11, 1: c13, 2: c33), against <g, dc>.
    Returns [(fd / <g, dc>) for each h]."""
    P = problem3()
    damping = orc.init_damping3d(P["nx"], P["ny"], P["nz"], P["npad"])
    padb, randb, _ = _shells(orc, P)
    obs = _obs3(P, damping)
    args = (P["souloc"], np.array([0]), P["recloc"], P["w"])
    rest = (P["nt"], P["npad"], P["o"], P["d"], P["dt"], P["rate"], obs)
    _, g = orc.gradient3d(orc.F64, *args, P["cij"].copy(), padb, randb, *rest)
    z, y, x = np.meshgrid(np.arange(P["nz"]), np.arange(P["ny"]), np.arange(P["nx"]), indexing="ij")
    dc = np.zeros_like(P["cij"])
    # integer-valued, so c +- h dc is exact in float for the h used
    dc[k] = np.round(2e4 * np.exp(-((z - 13) ** 2 + (y - 10) ** 2 + (x - 11) ** 2) / 8.0))
    gd = float(np.sum(g * dc))
    out = []
    for h in hs:
        fp, _ = orc.gradient3d(orc.F64, *args, (P["cij"] + h * dc).astype(np.float32), padb, randb, *rest)
        fm, _ = orc.gradient3d(orc.F64, *args, (P["cij"] - h * dc).astype(np.float32), padb, randb, *rest)
        out.append((fp - fm) / (2 * h) / gd)
    return out


# ---------------------------------------------------------------------------
# 2D RTM / ODCIG
# ---------------------------------------------------------------------------
def problem2(ratio, ntneg=0, seed=2):
    nx, nz, npad = 40, 32, 8
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    dx = dz = 10.0
    dt = 1e-3
    nt = 24 * ratio + 1 if ratio > 1 else 97
    rng = np.random.default_rng(seed)
    v = 2000.0 + 600.0 * rng.random((nnz, nnx))
    v[:, :] = np.maximum.accumulate(v, axis=0)  # roughly increasing with depth
    eps = 0.05 + 0.1 * rng.random((nnz, nnx))
    dl = 0.5 * eps
    c33 = (v * v).astype(np.float32)
    c11 = (c33 * (1 + 2 * eps)).astype(np.float32)
    c13 = (c33 * np.sqrt(1 + 2 * dl)).astype(np.float32)
    w = ricker(nt, dt, 30.0, 0.025)
    sloc = np.array([[20 + npad, 6 + npad, nx - 8, 0], [12 + npad, 5 + npad, nx - 8, nx - 8]], np.int32)
    rloc = []
    for _ in range(2):
        for ix in range(4, nx - 4):
            rloc.append((ix + npad, npad + 3))
    rloc = np.array(rloc, np.int32)
    rate = ratio * dt
    nnt_data = (nt - ntneg - 1) // ratio + 1
    data = (rng.standard_normal((nnt_data, len(rloc))) * 1e-3).astype(np.float32)
    return dict(nx=nx, nz=nz, npad=npad, nnx=nnx, nnz=nnz, dx=dx, dz=dz, dt=dt, nt=nt, rate=rate, ratio=ratio,
                ntneg=ntneg, c=(c11, c13, c33), w=w, sloc=sloc, rloc=rloc, data=data)


def taper2d(nx, nz, npad):
    """init_abc(taper, nx, nz, npad) (init.cpp:70-89) in numpy: an input of
    rtm_f / odcig_f here, not under test."""
    t0 = (0.95 + 0.05 * np.cos(np.pi * np.arange(npad - 1, -1, -1, dtype=np.float32).astype(np.float64) / npad))
    t0 = t0.astype(np.float32)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    tx = np.ones(nnx, np.float32)
    tx[:npad] = t0
    tx[npad + nx:] = t0[::-1]
    tz = np.ones(nnz, np.float32)
    tz[:npad] = t0
    tz[npad + nz:] = t0[::-1]
    return (tz[:, None] * tx[None, :]).astype(np.float32)


def rtm2d_history(orc, nhx=-1, ratio=4, ntneg=0):
    """Oracle FP64 rtm_f (nhx < 0) / odcig_f against the stored-history image."""
    P = problem2(ratio, ntneg)
    taper = taper2d(P["nx"], P["nz"], P["npad"])
    ot = -ntneg * P["dt"]
    img = orc.rtm2d(orc.F64, P["data"], *P["c"], P["w"], P["sloc"], P["rloc"], taper, P["nx"], P["nz"], P["nt"],
                    P["npad"], P["dx"], P["dz"], P["dt"], P["rate"], ot=ot, nhx=nhx)
    ib = bh.rtm2(P["data"], P["c"], P["w"], P["sloc"], P["rloc"], taper, P["nnx"], P["nnz"], P["nt"], P["dx"],
                 P["dz"], P["dt"], P["ratio"], ntneg, nhx)
    return rel(img, ib)


# ---------------------------------------------------------------------------
# every check, with the bar each must meet
# ---------------------------------------------------------------------------
def run_all(orc):
    """{name: (error, bar)} of every check above."""
    out = {}
    g, j = gradient3d_history(orc)
    out["gradient3d"] = (g, 1e-9)
    out["objective3d"] = (j, 1e-12)
    g, j = gradient3d_history(orc, tpow=1.5)
    out["gradient3d_tpow"] = (g, 1e-9)
    out["objective3d_tpow"] = (j, 1e-12)
    out["rtm3d"] = (rtm3d_history(orc), 1e-9)
    for nhx in (-1, 3):
        for ratio in (1, 4):
            out[f"rtm2d_nhx{nhx}_ratio{ratio}"] = (rtm2d_history(orc, nhx=nhx, ratio=ratio), 1e-9)
    out["odcig2d_ot_neg"] = (rtm2d_history(orc, nhx=2, ratio=2, ntneg=6), 1e-9)
    return out


if __name__ == "__main__":  # used by test_oracle_mutants.py in a child process
    import json
    import os
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import oracle

    print(json.dumps({k: [float(e), b] for k, (e, b) in run_all(oracle).items()}))
