"""Small seeded 3D VTI test problems shared by the 3D tests."""
import numpy as np


def layered_model(nz, ny, nx, seed=0, vmin=1500.0, vmax=4500.0, dip=0.15):
    """Seeded layered model with dipping interfaces: returns v, eps, del (nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    nl = 5
    z = np.arange(nz)[:, None, None].astype(np.float64)
    y = np.arange(ny)[None, :, None].astype(np.float64)
    x = np.arange(nx)[None, None, :].astype(np.float64)
    depth = z + dip * (x - nx / 2) + 0.5 * dip * (y - ny / 2)
    bounds = np.sort(rng.uniform(0.15, 0.9, nl - 1)) * nz
    layer = np.searchsorted(bounds, depth)
    vl = np.sort(rng.uniform(vmin, vmax, nl))
    el = rng.uniform(0.0, 0.25, nl)
    dl = np.minimum(rng.uniform(0.0, 0.12, nl), el)
    return vl[layer].astype(np.float32), el[layer].astype(np.float32), dl[layer].astype(np.float32)


def cij_from(v, eps, dl):
    c33 = (v.astype(np.float32) * v).astype(np.float32)
    c11 = (c33 * (1 + 2 * eps)).astype(np.float32)
    c13 = (c33 * np.sqrt(1 + 2 * dl)).astype(np.float32)
    return c11, c13, c33


def ricker(nt, dt, f, t0):
    t = np.arange(nt) * dt - t0
    a = (np.pi * f * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))
