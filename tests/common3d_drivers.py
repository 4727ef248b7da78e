"""A small 3D FWI/RTM problem with the reference's input conventions."""
import numpy as np

from common3d import cij_from, layered_model, ricker


def problem(nz=40, ny=36, nx=44, npad=8, nt=130, ns=2, seed=11, rate=4e-3):
    dx = 12.5
    o = (0.0, 0.0, 0.0)
    d = (dx, dx, dx)
    dt = 1e-3
    v, e, dl = layered_model(nz, ny, nx, seed=seed)
    c11, c13, c33 = cij_from(v, e, dl)
    cij = np.stack([c11, c13, c33]).astype(np.float32)
    # true model: a 4% faster blob in the middle
    vt = v.copy()
    vt[nz // 2 - 4: nz // 2 + 4, ny // 2 - 5: ny // 2 + 5, nx // 2 - 6: nx // 2 + 6] *= 1.04
    t11, t13, t33 = cij_from(vt, e, dl)
    ctrue = np.stack([t11, t13, t33]).astype(np.float32)
    w = ricker(nt, dt, 18.0, 0.06)
    # receivers: one line at depth index 9 (true z = 9*dx), every x interior, one y
    line = []
    for iy in (ny // 2 - 3, ny // 2 + 3):
        for ix in range(4, nx - 4):
            line.append((ix * dx + 0.1, iy * dx + 0.1, 9 * dx + 0.1, 1.0))
    nr = len(line)
    # each shot owns its own receiver range (souloc[4] = first receiver)
    recs = np.array(line * ns, np.float32)
    recs[5, 3] = 0.0  # one dead trace
    sou = []
    for k in range(ns):
        sou.append(((12 + 14 * k) * dx + 0.1, (ny // 2) * dx + 0.1, 12 * dx + 0.1, nr, k * nr))
    souloc = np.array(sou, np.float32)
    # random boundary for pass 2/3: scaled shell (deterministic, distinct)
    nbnd = nx * ny * nz - (nx - 2 * npad) * (ny - 2 * npad) * (nz - 2 * npad)
    rng = np.random.default_rng(seed + 1)
    shell_scale = rng.uniform(0.75, 1.0, nbnd).astype(np.float32)
    return dict(nx=nx, ny=ny, nz=nz, npad=npad, nt=nt, dt=dt, rate=rate, o=o, d=d, cij=cij, ctrue=ctrue, w=w,
                recloc=recs, souloc=souloc, shell_scale=shell_scale, nbnd=nbnd)
