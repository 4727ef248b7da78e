"""Seeded synthetic inputs of the benchmark configurations (models,
geometries, wavelets), in the reference's layouts, for bench.py and the
tests.  No data files: every array is generated from its seed.

- cfg2: 2D FWI on a 512 x 1024 grid (+40-point boundary), 10 m, 1 ms, 3000
  steps.  True model: the 1D gradient 1500 -> 4500 m/s with 64 seeded
  Gaussian anomalies (+-10 %, sigma 15 points); delta from 0 at the top to
  0.1 at the bottom, eps = 0.5 delta + 0.05.  Initial model: the true model
  smoothed with a sigma = 30 Gaussian.  32 sources (10 Hz Ricker) evenly
  spaced two rows below the surface, 1024 receivers one row below it.
- cfg3: the layered 3D model with dipping interfaces of the 460 x 572 x 572
  (30-point boundary) grid.
- cfg4: 256 shots, each on an aperture of 300-900 columns (seeded) of a
  4096-wide model built as cfg2's, nz 512, 2500 steps.
"""
from __future__ import annotations

import numpy as np


def ricker(nt, dt, freq, t0):
    t = np.arange(nt) * dt - t0
    a = (np.pi * freq * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _gauss_smooth(a, sigma):
    """Separable Gaussian smoothing (truncated at 4 sigma, edges repeated)."""
    r = int(4 * sigma)
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    out = a.astype(np.float64)
    for axis in range(out.ndim):
        pad = [(0, 0)] * out.ndim
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="edge")
        out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), axis, p)
    return out


def fwi2d_model(nx, nz, npad, seed, n_anom, smooth=30.0):
    """(true, initial) models as (v, eps, del) triples on the padded grid
    (nz + 2 npad, nx + 2 npad), float32."""
    rng = np.random.default_rng(seed)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    z = np.clip(np.arange(nnz) - npad, 0, nz - 1).astype(np.float64)
    v = np.repeat((1500.0 + 3000.0 * z / (nz - 1))[:, None], nnx, axis=1)
    zz, xx = np.meshgrid(np.arange(nnz), np.arange(nnx), indexing="ij")
    pert = np.zeros((nnz, nnx))
    for _ in range(n_anom):
        cz = rng.uniform(npad + 20, npad + nz - 20)
        cx = rng.uniform(npad + 20, npad + nx - 20)
        amp = rng.uniform(0.05, 0.10) * rng.choice([-1.0, 1.0])
        r2 = (zz - cz) ** 2 + (xx - cx) ** 2
        near = r2 < (5 * 15) ** 2
        pert[near] += amp * np.exp(-r2[near] / (2 * 15.0 ** 2))
    vt = v * (1 + np.clip(pert, -0.1, 0.1))
    dl = np.repeat((0.1 * z / (nz - 1))[:, None], nnx, axis=1)
    eps = 0.5 * dl + 0.05
    true = tuple(a.astype(np.float32) for a in (vt, eps, dl))
    init = tuple(_gauss_smooth(a, smooth).astype(np.float32) for a in (vt, eps, dl))
    return true, init


def cij(v, eps, dl):
    """c33 = v^2, c11 = c33 (1 + 2 eps), c13 = c33 sqrt(1 + 2 del) in float32."""
    v, eps, dl = (np.asarray(a, np.float32) for a in (v, eps, dl))
    c33 = (v * v).astype(np.float32)
    c11 = (c33 * (np.float32(1) + np.float32(2) * eps)).astype(np.float32)
    c13 = (c33 * np.sqrt(np.float32(1) + np.float32(2) * dl)).astype(np.float32)
    return c11, c13, c33


def cfg2(seed=2024, nshots=32):
    """The cfg2 problem: dict of sizes, true/initial models, geometry, wavelet."""
    nx, nz, npad, nt, dx, dt, rate = 1024, 512, 40, 3001, 10.0, 1e-3, 4e-3
    true, init = fwi2d_model(nx, nz, npad, seed, 64)
    sloc = np.array([[npad + 8 + (k * (nx - 16)) // max(nshots - 1, 1), npad + 2, nx, 0] for k in range(nshots)],
                    np.int32)
    rloc = np.array([[npad + i, npad + 1] for i in range(nx)], np.int32)
    return dict(nx=nx, nz=nz, npad=npad, nt=nt, dx=dx, dz=dx, dt=dt, rate=rate, true=true, init=init,
                w=ricker(nt, dt, 10.0, 0.12), sloc=sloc, rloc=rloc)


def layered3d(nz, ny, nx, seed=2024):
    """Seeded layered 3D model with dipping interfaces: (c11, c13, c33, (v, eps, del))."""
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
    return c11.astype(np.float32), c13.astype(np.float32), c33.astype(np.float32), (v, e, d)


def cfg4(seed=4, nshots=256, NX=4096, nz=512, npad=40, nt=2501):
    """The cfg4 batch: a NX-wide model (true and initial, as cfg2's with
    64 anomalies per 1024 columns), and `nshots` shots on apertures of
    300-900 columns: shots (c0, nx, ix, iz, nr, r0) and rloc (ix, iz) in each
    shot's sub-grid, a receiver in every column of the aperture."""
    rng = np.random.default_rng(seed)
    true, init = fwi2d_model(NX, nz, npad, seed, 64 * NX // 1024)
    shots, rloc = [], []
    for _ in range(nshots):
        nx = int(rng.integers(300, 901))
        c0 = int(rng.integers(0, NX - nx + 1))
        shots.append((c0, nx, npad + nx // 2, npad + 2, nx, len(rloc)))
        rloc += [(npad + i, npad + 1) for i in range(nx)]
    return dict(NX=NX, nz=nz, npad=npad, nt=nt, dx=10.0, dz=10.0, dt=1e-3, rate=4e-3, true=true, init=init,
                w=ricker(nt, 1e-3, 10.0, 0.12), shots=np.array(shots, np.int32), rloc=np.array(rloc, np.int32))
