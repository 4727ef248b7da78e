"""Oracle: CPU restatement of the reference's VTI path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The product (libvtiamd.so and its
binding) never loads it.

Parity status per function (what pins it) is documented in PINS below and
checked by tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_float, c_int, c_long, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VTI_ORACLE_LIB names another build of this library (the mutation test's
# deliberately broken builds); the default is the one built beside this file.
LIB_PATH = os.environ.get("VTI_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

F32, REF, F64 = 0, 1, 2

# What pins each oracle function (the checks live in tests/).
PINS = {
    "orc3d_init_damping": "bit-exact vs compiled reference abc3d.cu host code (tests/golden) + closed form",
    "orc3d_boundary": "bit-exact vs compiled reference boundary.cpp (tests/golden) + round trip",
    "orc3d_forward": "FP64 mode: plane-wave dispersion closed form, time reversibility, energy; "
    "REF/F32 modes within 1e-5 of FP64",
    "orc3d_gradient": "FP64 mode: equals a stored-history restatement (every source level kept, no "
    "random-boundary reversal) to 1e-14, with and without the t^p gain; Taylor ratios per coefficient; "
    "mutants of the receiver weights, gc13/gc33 terms, residual sign, t^p exponent and time index fail",
    "orc3d_rtm": "FP64 mode: equals the stored-history image to 1e-14 (image weights mutant fails); "
    "linear in the data",
    "orc2d_rtm": "FP64 mode: rtm_f and odcig_f equal a stored-history restatement (every level kept, "
    "closed-form slice and data interpolation) bit for bit, ratio 1/2/4, ot < 0; image, interpolation "
    "and offset mutants fail",
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built at {LIB_PATH}; run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        P = c_void_p
        L.orc3d_init_damping.argtypes = [P, c_int, c_int, c_int, c_int]
        L.orc3d_boundary.argtypes = [c_int, P, P, c_int, c_int, c_int, c_int]
        L.orc3d_nboundary.restype = c_long
        L.orc3d_nboundary.argtypes = [c_int] * 4
        L.orc3d_lmax.argtypes = [c_int, c_int]
        L.orc3d_forward.argtypes = [
            c_int, P, P, P, c_int, c_int, c_int, c_int, c_float, c_float, c_float, c_float, c_int, P, c_long,
            c_int, P, c_int, c_int, c_int, P, P, P, P, P, P,
        ]
        L.orc3d_steps.argtypes = [c_int, P, P, P, P, P, P, P] + [c_int] * 4 + [c_float] * 4 + [c_int, c_int]
        L.orc3d_steps_f64.argtypes = [P] * 7 + [c_int] * 4 + [c_float] * 4 + [c_int, c_int, c_int]
        L.orc3d_gradient.restype = c_double
        L.orc3d_gradient.argtypes = (
            [c_int, P, P, P, P, c_int, P, P, P, P, P]
            + [c_int] * 5
            + [c_float] * 9
            + [c_int, P, c_float, c_int, c_int]
        )
        L.orc3d_agc_residual.restype = c_double
        L.orc3d_agc_residual.argtypes = [c_int, P, P, c_int, c_int, c_int]
        L.orc3d_rtm.argtypes = [c_int, P, P, P, c_int, P, P, P, P, P] + [c_int] * 5 + [c_float] * 8 + [c_int, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def init_damping3d(nx, ny, nz, npad):
    d = np.zeros(nx * ny + nz, np.float32)
    lib().orc3d_init_damping(_p(d), nx, ny, nz, npad)
    return d


def lmax3d(nt, ngpu=1):
    return lib().orc3d_lmax(nt, ngpu)


def forward3d(mode, c11, c13, c33, npad, dx, dy, dz, dt, lmax, wavelet, sou_index, rec, sts, nnt,
              absorbing=True, want_state=False):
    """Returns (traces float32 (nr, nnt), traces float64 (nr, nnt), state or None).

    state = (cur_x, cur_z, prev_x, prev_z) as stored by the reference's
    block pipeline (cur damped once, prev damped twice in absorbing runs)."""
    nz, ny, nx = c11.shape
    c11, c13, c33 = (np.ascontiguousarray(a, np.float32) for a in (c11, c13, c33))
    w = np.ascontiguousarray(wavelet, np.float32)
    rec = np.ascontiguousarray(rec, np.int64)
    tr = np.zeros((rec.size, nnt), np.float32)
    tr64 = np.zeros((rec.size, nnt), np.float64)
    st = None
    sp = [None] * 4
    if want_state:
        dt_ = np.float64 if mode == F64 else np.float32
        st = [np.zeros((nz, ny, nx), dt_) for _ in range(4)]
        sp = [_p(a) for a in st]
    rc = lib().orc3d_forward(
        mode, _p(c11), _p(c13), _p(c33), nx, ny, nz, npad, dx, dy, dz, dt, lmax, _p(w), int(sou_index), rec.size,
        _p(rec), sts, nnt, int(absorbing), _p(tr), _p(tr64), *sp,
    )
    if rc != 0:
        raise RuntimeError(f"orc3d_forward failed {rc}")
    return tr, tr64, st


def steps3d(mode, state, c11, c13, c33, npad, dx, dy, dz, dt, nsteps, absorbing=True):
    """Advance (cur_x, prev_x, cur_z, prev_z) float32 arrays in place (mode F32 or REF)."""
    cx, px, cz, pz = state
    nz, ny, nx = c11.shape
    rc = lib().orc3d_steps(mode, _p(cx), _p(px), _p(cz), _p(pz), _p(c11), _p(c13), _p(c33), nx, ny, nz, npad,
                           dx, dy, dz, dt, nsteps, int(absorbing))
    if rc != 0:
        raise RuntimeError("orc3d_steps failed")


def bwd_steps3d(mode, arith, sigma, lam, acc, c11, c13, c33, npad, dx, dy, dz, dt, nsteps):
    """Backward levels with no source/receivers on float32 state, in place:
    sigma = [cur_x, prev_x, cur_z, prev_z], lam likewise, acc = [gc11, gc13,
    gc33] (mode 0) or [image] (mode 1).  Returns (sigma, lam) reordered so
    that each is again [cur_x, prev_x, cur_z, prev_z] after the level swaps."""
    L = lib()
    if not getattr(L, "_has_bwd", False):
        P = c_void_p
        L.orc3d_bwd_steps.argtypes = [c_int, c_int] + [P] * 14 + [c_int] * 4 + [c_float] * 4 + [c_int]
        L._has_bwd = True
    nz, ny, nx = c11.shape
    a = list(acc) + [None] * (3 - len(acc))
    rc = L.orc3d_bwd_steps(mode, arith, *[_p(x) for x in sigma], *[_p(x) for x in lam], *[_p(x) for x in a],
                           _p(c11), _p(c13), _p(c33), nx, ny, nz, npad, dx, dy, dz, dt, nsteps)
    if rc != 0:
        raise RuntimeError("orc3d_bwd_steps failed")
    if nsteps % 2:
        sigma = [sigma[1], sigma[0], sigma[3], sigma[2]]
        lam = [lam[1], lam[0], lam[3], lam[2]]
    return sigma, lam


def steps3d_f64(state, c11, c13, c33, npad, dx, dy, dz, dt, nsteps, absorbing=True, gradient_form=False):
    """Advance float64 (cur_x, prev_x, cur_z, prev_z) in place with the FP64 oracle."""
    cx, px, cz, pz = state
    nz, ny, nx = c11.shape
    c11, c13, c33 = (np.ascontiguousarray(a, np.float64) for a in (c11, c13, c33))
    rc = lib().orc3d_steps_f64(_p(cx), _p(px), _p(cz), _p(pz), _p(c11), _p(c13), _p(c33), nx, ny, nz, npad,
                               dx, dy, dz, dt, nsteps, int(absorbing), int(gradient_form))
    if rc != 0:
        raise RuntimeError("orc3d_steps_f64 failed")


def gradient3d(mode, souloc, shotid, recloc, wavelet, cij, padb, randb, nt, npad, o, d, dt, rate, obs, ngpu=1,
               tpow=0.0, objtype=0, halfwidth=20, ot=0.0):
    """objFuncGradientCij3d_f restated.  cij (3, nz, ny, nx) float32 modified in place.
    Returns (objective, gradient (3, nz, ny, nx)) -- float32 for F32/REF, float64 for F64."""
    _, nz, ny, nx = cij.shape
    souloc = np.ascontiguousarray(souloc, np.float32).ravel()
    recloc = np.ascontiguousarray(recloc, np.float32).ravel()
    shotid = np.ascontiguousarray(shotid, np.int32)
    w = np.ascontiguousarray(wavelet, np.float32)
    padb = np.ascontiguousarray(padb, np.float32)
    randb = np.ascontiguousarray(randb, np.float32)
    obs = np.ascontiguousarray(obs, np.float32)
    if mode == F64:
        out = np.zeros(1 + cij.size, np.float64)
        f = lib().orc3d_gradient(mode, None, _p(out), _p(souloc), _p(shotid), shotid.size, _p(recloc), _p(w), _p(cij),
                                 _p(padb), _p(randb), nx, ny, nz, nt, npad, *o, ot, *d, dt, rate, ngpu, _p(obs),
                                 tpow, objtype, halfwidth)
    else:
        out = np.zeros(1 + cij.size, np.float32)
        f = lib().orc3d_gradient(mode, _p(out), None, _p(souloc), _p(shotid), shotid.size, _p(recloc), _p(w), _p(cij),
                                 _p(padb), _p(randb), nx, ny, nz, nt, npad, *o, ot, *d, dt, rate, ngpu, _p(obs),
                                 tpow, objtype, halfwidth)
    return out[0], out[1:].reshape(3, nz, ny, nx)


def agc_residual(mode, d, d0, halfwidth):
    """residualAGC (agc.cpp:70-101) on trace-major (ntr, nt) traces: returns
    (sum of squared AGC residuals, adjoint source)."""
    ntr, nt = d.shape
    a = np.ascontiguousarray(d, np.float64 if mode == F64 else np.float32).copy()
    b = np.ascontiguousarray(d0, np.float32)
    fi = lib().orc3d_agc_residual(mode, _p(a), _p(b), nt, ntr, halfwidth)
    return fi, a


def rtm3d(mode, souloc, shotid, recloc3, wavelet, c11, c13, c33, nt, npad, o, d, dt, rate, obs, ngpu=1):
    nz, ny, nx = c11.shape
    souloc = np.ascontiguousarray(souloc, np.float32).ravel()
    recloc3 = np.ascontiguousarray(recloc3, np.float32).ravel()
    shotid = np.ascontiguousarray(shotid, np.int32)
    w = np.ascontiguousarray(wavelet, np.float32)
    c11, c13, c33 = (np.ascontiguousarray(a, np.float32) for a in (c11, c13, c33))
    obs = np.ascontiguousarray(obs, np.float32)
    img = np.zeros((nz, ny, nx), np.float64 if mode == F64 else np.float32)
    rc = lib().orc3d_rtm(mode, _p(img), _p(souloc), _p(shotid), shotid.size, _p(recloc3), _p(w), _p(c11), _p(c13),
                         _p(c33), nx, ny, nz, nt, npad, *o, *d, dt, rate, ngpu, _p(obs))
    if rc != 0:
        raise RuntimeError("orc3d_rtm failed")
    return img


def _lib2():
    L = lib()
    if not getattr(L, "_has2", False):
        P = c_void_p
        L.orc2d_synthetic.argtypes = [c_int, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5
        L.orc2d_modeling.argtypes = [c_int, P, P, P, P, P, c_int, P] + [c_int] * 4 + [c_float] * 5
        L.orc2d_fg.restype = c_double
        L.orc2d_fg.argtypes = [c_int, c_int, c_int, P, P, P, P, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 4 + \
            [c_float] * 5
        L.orc2d_rtm.argtypes = [c_int, c_int, c_int, P, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 5 + \
            [c_float] * 5
        L.orc2d_hessian.argtypes = [c_int, c_int, c_int, P, P, P, P, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 4 + \
            [c_float] * 5
        L.orc2d_forwardCD_f64.argtypes = [P] * 7 + [c_double] * 3 + [c_int, c_int]
        L.orc2d_backwardDC_f64.argtypes = [P] * 7 + [c_double] * 3 + [c_int, c_int]
        L._has2 = True
    return L


def synthetic2d(mode, c11, c13, c33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0):
    c11, c13, c33, wavelet, taper = (np.ascontiguousarray(a, np.float32).ravel() for a in (c11, c13, c33, wavelet, taper))
    sloc = np.ascontiguousarray(sloc, np.int32).ravel()
    rloc = np.ascontiguousarray(rloc, np.int32).ravel()
    ratio = int(rate / dt + 0.5)
    ntneg = int(round(abs(ot) / dt))
    nnt_data = (nt - ntneg - 1) // ratio + 1
    data = np.zeros((nnt_data, rloc.size // 2), np.float64 if mode == F64 else np.float32)
    rc = _lib2().orc2d_synthetic(mode, _p(data), _p(c11), _p(c13), _p(c33), _p(wavelet), _p(sloc), sloc.size // 4,
                                 _p(rloc), rloc.size // 2, _p(taper), nx, nz, nt, npad, dx, dz, dt, rate, ot)
    if rc != 0:
        raise RuntimeError("orc2d_synthetic failed")
    return data


def modeling2d(mode, c11, c13, c33, wavelet, slocxz, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0, init=None):
    """modeling_f (modeling_f.cu:10-86): p snapshots (nnt_data, nnz, nnx); slots
    it never records keep `init` (zeros by default)."""
    c11, c13, c33, wavelet, taper = (np.ascontiguousarray(a, np.float32).ravel() for a in (c11, c13, c33, wavelet, taper))
    q = float(np.float32(rate) / np.float32(dt))
    ratio = int(np.floor(q + 0.5))  # roundf: half away from zero
    ntneg = int(round(abs(ot) / dt))
    nnt_data = (nt - ntneg - 1) // ratio + 1
    shape = (nnt_data, nz + 2 * npad, nx + 2 * npad)
    dtype = np.float64 if mode == F64 else np.float32
    wf = np.zeros(shape, dtype) if init is None else np.array(init, dtype).reshape(shape)
    rc = _lib2().orc2d_modeling(mode, _p(wf), _p(c11), _p(c13), _p(c33), _p(wavelet), int(slocxz), _p(taper), nx, nz,
                                nt, npad, dx, dz, dt, rate, ot)
    if rc != 0:
        raise RuntimeError("orc2d_modeling failed")
    return wf


def fg2d(mode, born, data, c11, c13, c33, dc, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0,
         ngpu=1):
    """objFuncGradient_f (born=False) / GNhessian_f (born=True) restated literally,
    as the reference runs them on ngpu GPUs (shots dealt round robin; each GPU
    carries its buffers from one of its shots to the next).
    Returns (objective, gc (3, nnz, nnx))."""
    a = [np.ascontiguousarray(x, np.float32).ravel() for x in (c11, c13, c33, wavelet, taper)]
    dcs = [None, None, None] if dc is None else [np.ascontiguousarray(x, np.float32).ravel() for x in dc]
    data = None if data is None else np.ascontiguousarray(data, np.float32).ravel()
    sloc = np.ascontiguousarray(sloc, np.int32).ravel()
    rloc = np.ascontiguousarray(rloc, np.int32).ravel()
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    gc = np.zeros((3, nnz, nnx), np.float64 if mode == F64 else np.float32)
    f = _lib2().orc2d_fg(mode, int(born), int(ngpu), _p(gc), _p(data), _p(a[0]), _p(a[1]), _p(a[2]), *[_p(x) for x in dcs],
                         _p(a[3]), _p(sloc), sloc.size // 4, _p(rloc), rloc.size // 2, _p(a[4]), nx, nz, nt, npad, dx,
                         dz, dt, rate, ot)
    return f, gc


def hessian2d(mode, data, c11, c13, c33, dc, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0,
              ngpu=1, seed_bg=True):
    """hessian_f restated literally (its two snapshot rotations included), as
    the reference runs it on ngpu GPUs.  Returns H*dc as (3, nnz, nnx).
    seed_bg=False fills the Born snapshot pipeline's first buffer from the
    Born snapshots instead of the background ones (a check only)."""
    a = [np.ascontiguousarray(x, np.float32).ravel() for x in (c11, c13, c33, wavelet, taper, data)]
    dcs = [np.ascontiguousarray(x, np.float32).ravel() for x in dc]
    sloc = np.ascontiguousarray(sloc, np.int32).ravel()
    rloc = np.ascontiguousarray(rloc, np.int32).ravel()
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    gc = np.zeros((3, nnz, nnx), np.float64 if mode == F64 else np.float32)
    rc = _lib2().orc2d_hessian(mode, int(seed_bg), int(ngpu), _p(gc), _p(a[5]), _p(a[0]), _p(a[1]), _p(a[2]), *[_p(x) for x in dcs],
                               _p(a[3]), _p(sloc), sloc.size // 4, _p(rloc), rloc.size // 2, _p(a[4]), nx, nz, nt, npad,
                               dx, dz, dt, rate, ot)
    if rc != 0:
        raise RuntimeError("orc2d_hessian failed")
    return gc


def rtm2d(mode, data, c11, c13, c33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0,
          nhx=-1, ngpu=1):
    """rtm_f (nhx < 0) or odcig_f (nhx >= 0) restated literally, as the
    reference runs them on ngpu GPUs.  Returns the image (nnz, nnx) or the
    offset gathers (2*nhx+1, nnz, nnx)."""
    c11, c13, c33, wavelet, taper, data = (np.ascontiguousarray(a, np.float32).ravel()
                                           for a in (c11, c13, c33, wavelet, taper, data))
    sloc = np.ascontiguousarray(sloc, np.int32).ravel()
    rloc = np.ascontiguousarray(rloc, np.int32).ravel()
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    shape = (nnz, nnx) if nhx < 0 else (2 * nhx + 1, nnz, nnx)
    img = np.zeros(shape, np.float64 if mode == F64 else np.float32)
    rc = _lib2().orc2d_rtm(mode, int(nhx >= 0), int(ngpu), _p(img), _p(data), _p(c11), _p(c13), _p(c33), _p(wavelet),
                           _p(sloc), sloc.size // 4, _p(rloc), rloc.size // 2, _p(taper), nx, nz, max(nhx, 0), nt, npad,
                           dx, dz, dt, rate, ot)
    if rc != 0:
        raise RuntimeError("orc2d_rtm failed")
    return img


def cpu3d_steps(mode, state, a, b, c, npad, dx, dy, dz, dt, nsteps, src_index=-1, wavelet=None):
    """The reference CPU path (laplacian3d.ispc forward/forwardCij + the
    modeling loop of vtifd3d8o-modeling.cpp): advance state = [prev_x, cur_x,
    prev_z, cur_z] (float32, (nz, ny, nx)) in place by nsteps levels.
    mode 0: a, b, c = v, eps, del; mode 1: c11, c13, c33."""
    L = lib()
    if not getattr(L, "_has_cpu", False):
        P = c_void_p
        L.cpu3d_steps.argtypes = [c_int] + [P] * 7 + [c_int] * 4 + [c_float] * 4 + [c_int, c_long, P]
        L.cpu3d_damping.argtypes = [P, c_int]
        L._has_cpu = True
    nz, ny, nx = a.shape
    a, b, c = (np.ascontiguousarray(x, np.float32) for x in (a, b, c))
    w = None if wavelet is None else np.ascontiguousarray(wavelet, np.float32)
    rc = L.cpu3d_steps(mode, *[_p(s) for s in state], _p(a), _p(b), _p(c), nx, ny, nz, npad, dx, dy, dz, dt, nsteps,
                       int(src_index), _p(w))
    if rc != 0:
        raise RuntimeError("cpu3d_steps failed")


def cpu3d_backward_steps(sigma, lam, gc, v, eps, dl, npad, dx, dy, dz, dt, nsteps):
    """The reference CPU gradient's backward level (gradientCij + forward on
    the adjoint + abc) on float32 state, in place: sigma = [prev_x, cur_x,
    prev_z, cur_z], lam likewise, gc = [gc11, gc13, gc33]; no injections."""
    L = lib()
    if not getattr(L, "_has_cpub", False):
        P = c_void_p
        L.cpu3d_backward_steps.argtypes = [P] * 14 + [c_int] * 4 + [c_float] * 4 + [c_int]
        L._has_cpub = True
    nz, ny, nx = v.shape
    v, eps, dl = (np.ascontiguousarray(x, np.float32) for x in (v, eps, dl))
    rc = L.cpu3d_backward_steps(*[_p(s) for s in sigma], *[_p(s) for s in lam], *[_p(g) for g in gc], _p(v), _p(eps),
                                _p(dl), nx, ny, nz, npad, dx, dy, dz, dt, nsteps)
    if rc != 0:
        raise RuntimeError("cpu3d_backward_steps failed")


def cpu3d_damping(npad):
    lib()
    cpu3d_steps  # binds argtypes lazily
    L = lib()
    L.cpu3d_damping.argtypes = [c_void_p, c_int]
    d = np.zeros(npad, np.float32)
    L.cpu3d_damping(_p(d), npad)
    return d
