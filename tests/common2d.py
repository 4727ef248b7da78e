"""Small seeded 2D VTI problems with the reference's input conventions.

Models live on the padded grid (nnz, nnx), x fastest; sloc holds
{ix, iz, nr_shot, first receiver} per shot and rloc {ix, iz} per receiver,
both in padded grid indices, as lib/vti-fwi-2d/gradient-param.cu:30-48
builds them."""
import numpy as np

from common3d import cij_from, ricker


def layered2d(nnz, nnx, seed=0, vmin=1500.0, vmax=3000.0):
    rng = np.random.default_rng(seed)
    z = np.arange(nnz)[:, None].astype(np.float64)
    x = np.arange(nnx)[None, :].astype(np.float64)
    depth = z + 0.12 * (x - nnx / 2)
    bounds = np.sort(rng.uniform(0.2, 0.85, 3)) * nnz
    layer = np.searchsorted(bounds, depth)
    vl = np.sort(rng.uniform(vmin, vmax, 4))
    # eps > del strictly: c11*c33 - c13^2 = 2 v^4 (eps - del) stays positive
    # under the perturbations the Taylor tests apply
    el = rng.uniform(0.08, 0.2, 4)
    dl = el * rng.uniform(0.2, 0.6, 4)
    return vl[layer].astype(np.float32), el[layer].astype(np.float32), dl[layer].astype(np.float32)


def problem2d(vti, nx=64, nz=48, npad=12, nt=201, ns=2, seed=3, rate=4e-3, dt=1e-3, dx=10.0, rec_step=2,
              blob=1.05):
    """Returns a dict with the background model (c), a perturbed true model
    (ctrue, a faster blob), wavelet, sloc/rloc and the taper."""
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    v, e, dl = layered2d(nnz, nnx, seed)
    c = cij_from(v, e, dl)
    vt = v.copy()
    vt[nnz // 2 - 5: nnz // 2 + 5, nnx // 2 - 8: nnx // 2 + 8] *= blob
    ctrue = cij_from(vt, e, dl)
    w = ricker(nt, dt, 20.0, 0.05)
    rx = list(range(npad, npad + nx, rec_step))
    rloc, sloc = [], []
    for k in range(ns):
        sloc.append((npad + (k + 1) * nx // (ns + 1), npad + 3, len(rx), len(rloc)))
        rloc += [(x, npad + 1) for x in rx]
    taper = vti.init_taper2d(nx, nz, npad).reshape(nnz, nnx)
    return dict(nx=nx, nz=nz, npad=npad, nnx=nnx, nnz=nnz, nt=nt, dt=dt, dx=dx, dz=dx, rate=rate, c=c,
                ctrue=ctrue, w=w, sloc=np.array(sloc, np.int32), rloc=np.array(rloc, np.int32), taper=taper,
                nr=len(rloc))


def args2d(P):
    """Trailing scalar arguments (nx, nz, nt, npad, dx, dz, dt, rate)."""
    return (P["nx"], P["nz"], P["nt"], P["npad"], P["dx"], P["dz"], P["dt"], P["rate"])
