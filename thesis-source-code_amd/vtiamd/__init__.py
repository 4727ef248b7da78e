"""ctypes binding of libvtiamd.so (the C ABI declared in include/vti_amd.h).

This is the host-side mirror used by tests, the benchmark and Python callers.
It loads the in-tree library and raises if it is missing: there is no
fallback path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_longlong, c_size_t, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libvtiamd.so")


class VtiError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise VtiError(f"libvtiamd.so not built at {LIB_PATH}; run __graft_entry__.build() first")
    return ctypes.CDLL(LIB_PATH)


_lib = _load()
P = c_void_p

_SIGS = {
    "vti_last_error": (c_char_p, []),
    "vti_ctx_create": (c_int, [POINTER(c_void_p), c_int]),
    "vti_ctx_destroy": (c_int, [P]),
    "vti_ctx_synchronize": (c_int, [P]),
    "vti_ctx_set_arith": (c_int, [P, c_int]),
    "vti_ctx_set_snapshots": (c_int, [P, c_int]),
    "vti_init_damping3d": (c_int, [P, c_int, c_int, c_int, c_int]),
    "vti_init_taper2d": (c_int, [P, c_int, c_int, c_int]),
    "vti_vepsdel_to_cij": (c_int, [P, P, P, P, P, P, c_float, c_float, c_float, c_longlong]),
    "vti_dvepsdel_to_dcij": (c_int, [P] * 12 + [c_float, c_float, c_float, c_longlong]),
    "vti_gradcij_to_gradvepsdel": (c_int, [P] * 9 + [c_float, c_float, c_float, c_longlong]),
    "vti_boundary_size": (c_longlong, [c_int, c_int, c_int, c_int]),
    "vti_get_boundary": (c_int, [P, P, c_int, c_int, c_int, c_int]),
    "vti_put_boundary": (c_int, [P, P, c_int, c_int, c_int, c_int]),
    "vti_zero_boundary": (c_int, [P, c_int, c_int, c_int, c_int]),
    "vti_ricker": (c_int, [P, c_float, c_int, c_float, c_float, c_float, c_float]),
    "vti_glibc_rand": (c_int, [c_int, c_int, P]),
    "vti_random_boundary": (
        c_longlong,
        [P, c_float, c_float, c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_int],
    ),
    "vti3d_prop_create": (
        c_int,
        [P, POINTER(c_void_p), c_int, c_int, c_int, c_int, c_float, c_float, c_float, c_float],
    ),
    "vti3d_prop_destroy": (c_int, [P]),
    "vti3d_prop_bytes": (c_size_t, [P]),
    "vti3d_prop_set_zchunk": (c_int, [P, c_int]),
    "vti3d_prop_set_levels": (c_int, [P, c_int]),
    "vti3d_prop_set_arith": (c_int, [P, c_int]),
    "vti3d_prop_set_model": (c_int, [P, P, P, P]),
    "vti3d_prop_reset": (c_int, [P, c_longlong, c_float]),
    "vti3d_prop_set_state": (c_int, [P, P, P, P, P]),
    "vti3d_prop_get_state": (c_int, [P, P, P, P, P]),
    "vti3d_prop_forward_steps": (c_int, [P, c_int, c_int, POINTER(c_float)]),
    "vti3d_prop_forward_shot": (
        c_int,
        [P, c_int, P, c_longlong, c_int, P, c_int, c_int, c_int, P],
    ),
    "vti3d_prop_backward_steps": (c_int, [P, c_int, c_int, POINTER(c_float)]),
    "vti3d_prop_set_adjoint": (c_int, [P, c_int, P, P, P, P]),
    "vti3d_prop_get_adjoint": (c_int, [P, P, P, P, P, P, P, P]),
    "vti3d_model_data": (
        c_int,
        [P, P, P, c_int, P, c_int, P, P, P, P, P] + [c_int] * 5 + [c_float] * 9 + [c_int],
    ),
    "vti3d_objfunc": (
        c_int,
        [P, P, P, c_int, P, c_int, P, P, P, P] + [c_int] * 5 + [c_float] * 9 + [c_int, P, c_float, c_int, c_int],
    ),
    "vti3d_objfunc_gradient": (
        c_int,
        [P, P, P, c_int, P, c_int, P, P, P, P, P] + [c_int] * 5 + [c_float] * 9 + [c_int, P, c_float, c_int, c_int],
    ),
    "vti3d_rtm": (
        c_int,
        [P, P, P, c_int, P, c_int, P, P, P, P, P] + [c_int] * 5 + [c_float] * 9 + [c_int, P],
    ),
    "vti2d_synthetic": (c_int, [P, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5),
    "vti2d_synthetic_noalloc": (c_int, [P, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5),
    "vti2d_modeling": (c_int, [P, P, P, P, P, P, c_int, P] + [c_int] * 4 + [c_float] * 5),
    "vti2d_hessian": (c_int, [P] * 13 + [c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5),
    "vti2d_rtm": (c_int, [P] * 7 + [P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5),
    "vti2d_odcig": (c_int, [P] * 7 + [P, c_int, P, c_int, P] + [c_int] * 5 + [c_float] * 5),
    "vti2d_objfunc_gradient": (
        c_int,
        [P, P, P, P, P, P, P, P, P, P, P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5,
    ),
    "vti2d_gn_hessian": (
        c_int,
        [P] * 13 + [c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 5,
    ),
    "vti_check_eps_del": (c_int, [P, P, c_float, c_float, c_longlong, P]),
    "vti2d_batch_tune": (c_int, [P, P, c_int, P, c_int] + [c_int] * 4 + [c_float] * 4
                         + [POINTER(c_int)] * 3 + [POINTER(ctypes.c_double)]),
    "vti2d_objfunc_gradient_batch": (
        c_int,
        [P] * 11 + [c_int, P] + [c_int] * 5 + [c_float] * 4 + [c_int] * 4,
    ),
    "vti2d_batch_plan_create": (c_int, [P, POINTER(c_void_p), P, P, P, c_int, P, c_int] + [c_int] * 4
                                + [c_float] * 4 + [c_int] * 4),
    "vti2d_batch_plan_run": (c_int, [P] * 8),
    "vti2d_batch_plan_destroy": (c_int, [P]),
    "vti_lbfgs_create": (c_int, [POINTER(c_void_p), c_longlong, c_int, c_int]),
    "vti_lbfgs_create_device": (c_int, [POINTER(c_void_p), P, c_longlong, c_int, c_int]),
    "vti_lbfgs_destroy": (c_int, [P]),
    "vti_lbfgs_step": (c_int, [P, P, ctypes.c_double, P, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "vti_serve_create": (c_int, [P, c_int, c_int, ctypes.c_double]),
    "vti_serve_destroy": (c_int, [P]),
    "vti_serve_put_model": (c_int, [P, P, P, P, c_longlong, POINTER(c_int)]),
    "vti_serve_drop_model": (c_int, [P, c_int]),
    "vti_serve_grad2d": (c_int, [P, POINTER(c_longlong), c_int] + [P] * 7 + [c_int, P, c_int, P] + [c_int] * 4
                         + [c_float] * 4),
    "vti_serve_model3d": (c_int, [P, POINTER(c_longlong), c_int, P, P, c_int, P, c_int, P, c_int, P] + [c_int] * 5
                          + [c_float] * 9),
    "vti_serve_wait": (c_int, [P, c_longlong, POINTER(ctypes.c_double), POINTER(ctypes.c_double)]),
    "vti_serve_stats": (c_int, [P] + [POINTER(ctypes.c_double)] * 3 + [POINTER(c_longlong), POINTER(ctypes.c_double)]),
    "vti_serve_set_batching": (c_int, [P, c_int]),
    "vti_serve_set_snapshots": (c_int, [P, c_int]),
    "vti_serve_set_exclusive3d": (c_int, [P, c_int]),
    "vti_serve_batch_stats": (c_int, [P, POINTER(c_longlong), POINTER(c_longlong)]),
    "vti2d_objfunc_gradient_vepsdel": (
        c_int,
        [P] * 9 + [c_float] * 3 + [P, P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 6 + [P],
    ),
    "vti2d_gn_hessian_vepsdel": (
        c_int,
        [P] * 11 + [c_float] * 3 + [P, P, c_int, P, c_int, P] + [c_int] * 4 + [c_float] * 6 + [P],
    ),
}

EXPORTED = tuple(_SIGS)

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

lib = _lib


def last_error() -> str:
    msg = _lib.vti_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        raise VtiError(f"{what} failed ({rc}): {last_error()}")
    return rc


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


# ---- host helpers ---------------------------------------------------------

def init_damping3d(nx, ny, nz, npad) -> np.ndarray:
    d = np.zeros(nx * ny + nz, np.float32)
    check(_lib.vti_init_damping3d(ptr(d), nx, ny, nz, npad), "vti_init_damping3d")
    return d


def init_taper2d(nx, nz, npad) -> np.ndarray:
    t = np.zeros((nz + 2 * npad) * (nx + 2 * npad), np.float32)
    check(_lib.vti_init_taper2d(ptr(t), nx, nz, npad), "vti_init_taper2d")
    return t


def vepsdel_to_cij(v, eps, dl, v0=1.0, eps0=1.0, del0=1.0):
    v, eps, dl = f32(v), f32(eps), f32(dl)
    c11, c13, c33 = (np.empty_like(v) for _ in range(3))
    check(_lib.vti_vepsdel_to_cij(ptr(c11), ptr(c13), ptr(c33), ptr(v), ptr(eps), ptr(dl), v0, eps0, del0, v.size))
    return c11, c13, c33


def glibc_rand(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.int32)
    check(_lib.vti_glibc_rand(seed, n, ptr(out)), "vti_glibc_rand")
    return out


def random_boundary(field, bval, eval_, dims, thick, pctG, seed) -> np.ndarray:
    """RandomMY on a copy of `field` (dims = (n1, n2, n3), n1 fastest)."""
    f = f32(field).copy()
    n1, n2, n3 = dims
    t1, t2, t3 = thick
    check(_lib.vti_random_boundary(ptr(f), bval, eval_, n1, n2, n3, t1, t2, t3, pctG, seed), "vti_random_boundary")
    return f


def ricker(freq, nt, dt, ot=0.0, tdelay=0.0, scalefactor=1.0) -> np.ndarray:
    w = np.zeros(nt, np.float32)
    check(_lib.vti_ricker(ptr(w), freq, nt, dt, ot, tdelay, scalefactor), "vti_ricker")
    return w


# ---- device objects ---------------------------------------------------------

class Context:
    def __init__(self, device: int = 0):
        h = c_void_p()
        check(_lib.vti_ctx_create(ctypes.byref(h), device), "vti_ctx_create")
        self.h = h

    def sync(self):
        check(_lib.vti_ctx_synchronize(self.h), "vti_ctx_synchronize")

    def set_arith(self, arith: int):
        """1: the reference's mixed FP64 stencil arithmetic (default); 0: FP32."""
        check(_lib.vti_ctx_set_arith(self.h, arith), "vti_ctx_set_arith")

    def set_snapshots(self, snap: str):
        """Snapshot storage of 2D gradients on this context: "fp32" (default),
        "fp16", "bf16" or "s16"; 16-bit forms apply to calls the batched
        driver takes (see vti_ctx_set_snapshots)."""
        check(_lib.vti_ctx_set_snapshots(self.h, SNAP_FORMATS[snap]), "vti_ctx_set_snapshots")

    def close(self):
        if self.h:
            check(_lib.vti_ctx_destroy(self.h), "vti_ctx_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Prop3D:
    """Device-resident 3D VTI propagator (vti3d_prop_*)."""

    def __init__(self, ctx: Context, nx, ny, nz, npad, dx, dy, dz, dt):
        h = c_void_p()
        check(_lib.vti3d_prop_create(ctx.h, ctypes.byref(h), nx, ny, nz, npad, dx, dy, dz, dt), "vti3d_prop_create")
        self.h = h
        self.shape = (nz, ny, nx)

    def close(self):
        if self.h:
            check(_lib.vti3d_prop_destroy(self.h), "vti3d_prop_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nbytes(self) -> int:
        return int(_lib.vti3d_prop_bytes(self.h))

    def set_zchunk(self, zc: int):
        check(_lib.vti3d_prop_set_zchunk(self.h, zc), "vti3d_prop_set_zchunk")

    def set_levels(self, levels):
        """Forward levels per launch of forward_steps: 1 or 2 (same bits)."""
        check(_lib.vti3d_prop_set_levels(self.h, levels), "vti3d_prop_set_levels")

    def set_arith(self, arith: int):
        check(_lib.vti3d_prop_set_arith(self.h, arith), "vti3d_prop_set_arith")

    def set_model(self, c11, c13, c33):
        c11, c13, c33 = f32(c11), f32(c13), f32(c33)
        check(_lib.vti3d_prop_set_model(self.h, ptr(c11), ptr(c13), ptr(c33)), "vti3d_prop_set_model")

    def reset(self, sou_index: int = -1, value: float = 0.0):
        check(_lib.vti3d_prop_reset(self.h, sou_index, value), "vti3d_prop_reset")

    def set_state(self, cur_x=None, cur_z=None, prev_x=None, prev_z=None):
        arrs = [None if a is None else f32(a) for a in (cur_x, cur_z, prev_x, prev_z)]
        check(_lib.vti3d_prop_set_state(self.h, *[ptr(a) for a in arrs]), "vti3d_prop_set_state")

    def get_state(self):
        outs = [np.empty(self.shape, np.float32) for _ in range(4)]
        check(_lib.vti3d_prop_get_state(self.h, *[ptr(a) for a in outs]), "vti3d_prop_get_state")
        return outs

    def forward_steps(self, nsteps: int, absorbing: bool = True, timed: bool = False):
        ms = c_float(0.0)
        check(
            _lib.vti3d_prop_forward_steps(self.h, nsteps, int(absorbing), ctypes.byref(ms) if timed else None),
            "vti3d_prop_forward_steps",
        )
        return ms.value if timed else None

    def forward_shot(self, lmax, wavelet, sou_index, rec, sts, nnt, absorbing=True):
        """Returns traces (nr, nnt), trace-major like the reference; sample 0 = 0."""
        wavelet = f32(wavelet)
        rec = np.ascontiguousarray(rec, dtype=np.int64)
        tr = np.zeros((nnt, rec.size), np.float32)
        check(
            _lib.vti3d_prop_forward_shot(
                self.h, lmax, ptr(wavelet), int(sou_index), rec.size, ptr(rec), sts, nnt, int(absorbing), ptr(tr)
            ),
            "vti3d_prop_forward_shot",
        )
        return np.ascontiguousarray(tr.T)

    def set_adjoint(self, mode, cur_x=None, cur_z=None, prev_x=None, prev_z=None):
        """Load lambda (reference layout) and clear the accumulators of `mode`."""
        a = [None if x is None else f32(x) for x in (cur_x, cur_z, prev_x, prev_z)]
        check(_lib.vti3d_prop_set_adjoint(self.h, mode, *[ptr(x) for x in a]), "vti3d_prop_set_adjoint")

    def get_adjoint(self, mode):
        """(lambda cur_x, cur_z, prev_x, prev_z, accumulators): the
        accumulators are (gc11, gc13, gc33) for mode 0, (image,) for mode 1."""
        out = [np.zeros(self.shape, np.float32) for _ in range(4 + (3 if mode == 0 else 1))]
        acc = out[4:] + [None] * (3 - len(out[4:]))
        check(_lib.vti3d_prop_get_adjoint(self.h, *[ptr(x) for x in out[:4]], *[ptr(x) for x in acc]),
              "vti3d_prop_get_adjoint")
        return out[0], out[1], out[2], out[3], tuple(out[4:])

    def backward_steps(self, nsteps: int, mode: int = 0, timed: bool = False):
        ms = c_float(0.0)
        check(
            _lib.vti3d_prop_backward_steps(self.h, mode, nsteps, ctypes.byref(ms) if timed else None),
            "vti3d_prop_backward_steps",
        )
        return ms.value if timed else None


# ---- 3D drivers (reference semantics) ----------------------------------------

def _geom3(souloc, shotid, recloc, wavelet):
    souloc = f32(souloc).ravel()
    recloc = f32(recloc).ravel()
    shotid = np.ascontiguousarray(shotid, dtype=np.int32)
    wavelet = f32(wavelet)
    return souloc, shotid, recloc, wavelet


def model_data3d(ctx, souloc, shotid, recloc, wavelet, c11, c13, c33, nt, npad, o, d, dt, rate, ngpu=1):
    """modelDataCij3d_f: returns the trace-major data array (nr_total * nnt)."""
    souloc, shotid, recloc, wavelet = _geom3(souloc, shotid, recloc, wavelet)
    c11, c13, c33 = f32(c11), f32(c13), f32(c33)
    nz, ny, nx = c11.shape
    sts = int(np.round(np.float32(rate) / np.float32(dt)))
    nnt = (nt - 1) // sts + 1
    nrt = recloc.size // 4
    dd = np.zeros(nrt * nnt, np.float32)
    check(_lib.vti3d_model_data(ctx.h, ptr(dd), ptr(souloc), souloc.size // 5, ptr(shotid), shotid.size, ptr(recloc),
                                ptr(wavelet), ptr(c11), ptr(c13), ptr(c33), nx, ny, nz, nt, npad, *o, 0.0, *d, dt,
                                rate, ngpu), "vti3d_model_data")
    return dd


def objfunc_gradient3d(ctx, souloc, shotid, recloc, wavelet, cij, padb, randb, nt, npad, o, d, dt, rate, obs,
                       ngpu=1, tpow=0.0, gradient=True, objtype=0, halfwidth=20, ot=0.0):
    """objFuncGradientCij3d_f / objFuncCij3d_f.  cij (3, nz, ny, nx) is modified in place
    like the reference (its boundary shell ends up holding randb / padb)."""
    souloc, shotid, recloc, wavelet = _geom3(souloc, shotid, recloc, wavelet)
    assert cij.dtype == np.float32 and cij.flags["C_CONTIGUOUS"]
    _, nz, ny, nx = cij.shape
    padb, obs = f32(padb), f32(obs)
    if gradient:
        randb = f32(randb)
        out = np.zeros(1 + cij.size, np.float32)
        check(_lib.vti3d_objfunc_gradient(ctx.h, ptr(out), ptr(souloc), souloc.size // 5, ptr(shotid), shotid.size,
                                          ptr(recloc), ptr(wavelet), ptr(cij), ptr(padb), ptr(randb), nx, ny, nz, nt,
                                          npad, *o, ot, *d, dt, rate, ngpu, ptr(obs), tpow, objtype, halfwidth),
              "vti3d_objfunc_gradient")
        return float(out[0]), out[1:].reshape(3, nz, ny, nx)
    out = np.zeros(1, np.float32)
    check(_lib.vti3d_objfunc(ctx.h, ptr(out), ptr(souloc), souloc.size // 5, ptr(shotid), shotid.size, ptr(recloc),
                             ptr(wavelet), ptr(cij), ptr(padb), nx, ny, nz, nt, npad, *o, ot, *d, dt, rate, ngpu,
                             ptr(obs), tpow, objtype, halfwidth), "vti3d_objfunc")
    return float(out[0])


def rtm3d(ctx, souloc, shotid, recloc3, wavelet, c11, c13, c33, nt, npad, o, d, dt, rate, obs, ngpu=1):
    """rtmCij3d_f: returns the image (nz, ny, nx).  recloc3 has 3 floats per receiver."""
    souloc, shotid, recloc3, wavelet = _geom3(souloc, shotid, recloc3, wavelet)
    c11, c13, c33, obs = f32(c11), f32(c13), f32(c33), f32(obs)
    nz, ny, nx = c11.shape
    img = np.zeros((nz, ny, nx), np.float32)
    check(_lib.vti3d_rtm(ctx.h, ptr(img), ptr(souloc), souloc.size // 5, ptr(shotid), shotid.size, ptr(recloc3),
                         ptr(wavelet), ptr(c11), ptr(c13), ptr(c33), nx, ny, nz, nt, npad, *o, 0.0, *d, dt, rate,
                         ngpu, ptr(obs)), "vti3d_rtm")
    return img


# ---- 2D drivers (reference semantics) ---------------------------------------

def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32).ravel()


def synthetic2d(ctx, c11, c13, c33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0,
                data=None, noalloc=False):
    """synthetic_f: returns data (nnt_data, nr) time-major.  Rows the
    reference never records keep the values of `data` (written in place when
    given as a C-contiguous float32 array of that shape), else zeros.
    noalloc: call vti2d_synthetic_noalloc (no device allocation; grids up to
    2^20 points)."""
    c11, c13, c33, wavelet, taper = (f32(a).ravel() for a in (c11, c13, c33, wavelet, taper))
    sloc, rloc = _i32(sloc), _i32(rloc)
    ns, nr = sloc.size // 4, rloc.size // 2
    ratio = int(rate / dt + 0.5)
    ntneg = int(round(abs(ot) / dt))
    nnt_data = (nt - ntneg - 1) // ratio + 1
    if data is None:
        data = np.zeros((nnt_data, nr), np.float32)
    elif data.dtype != np.float32 or data.shape != (nnt_data, nr) or not data.flags["C_CONTIGUOUS"]:
        raise ValueError(f"data must be a C-contiguous float32 array of shape {(nnt_data, nr)}")
    name = "vti2d_synthetic_noalloc" if noalloc else "vti2d_synthetic"
    check(getattr(_lib, name)(ctx.h, ptr(data), ptr(c11), ptr(c13), ptr(c33), ptr(wavelet), ptr(sloc), ns, ptr(rloc),
                              nr, ptr(taper), nx, nz, nt, npad, dx, dz, dt, rate, ot), name)
    return data


def modeling2d(ctx, c11, c13, c33, wavelet, slocxz, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0,
               wavefield=None):
    """modeling_f: returns the p snapshots (nnt_data, nnz, nnx).  Slots the
    reference never records keep the values of `wavefield` (written in place
    when given as a C-contiguous float32 array of that shape), else zeros."""
    c11, c13, c33, wavelet, taper = (f32(a).ravel() for a in (c11, c13, c33, wavelet, taper))
    q = float(np.float32(rate) / np.float32(dt))
    ratio = int(np.floor(q + 0.5))  # round(rate/dt) in float, half away from zero
    ntneg = int(round(abs(ot) / dt))
    shape = ((nt - ntneg - 1) // ratio + 1, nz + 2 * npad, nx + 2 * npad)
    if wavefield is None:
        wavefield = np.zeros(shape, np.float32)
    elif wavefield.dtype != np.float32 or wavefield.shape != shape or not wavefield.flags["C_CONTIGUOUS"]:
        raise ValueError(f"wavefield must be a C-contiguous float32 array of shape {shape}")
    check(_lib.vti2d_modeling(ctx.h, ptr(wavefield), ptr(c11), ptr(c13), ptr(c33), ptr(wavelet), int(slocxz), ptr(taper),
                              nx, nz, nt, npad, dx, dz, dt, rate, ot), "vti2d_modeling")
    return wavefield


def objfunc_gradient2d(ctx, data, c11, c13, c33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate,
                       ot=0.0):
    """objFuncGradient_f: returns (objective, gc (3, nnz, nnx))."""
    c11, c13, c33, wavelet, taper, data = (f32(a).ravel() for a in (c11, c13, c33, wavelet, taper, data))
    sloc, rloc = _i32(sloc), _i32(rloc)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    gc = np.zeros((3, nnz, nnx), np.float32)
    f = ctypes.c_float(0.0)
    check(_lib.vti2d_objfunc_gradient(ctx.h, ctypes.byref(f), ptr(gc[0]), ptr(gc[1]), ptr(gc[2]), ptr(data), ptr(c11),
                                      ptr(c13), ptr(c33), ptr(wavelet), ptr(sloc), sloc.size // 4, ptr(rloc),
                                      rloc.size // 2, ptr(taper), nx, nz, nt, npad, dx, dz, dt, rate, ot),
          "vti2d_objfunc_gradient")
    return f.value, gc


def hessian2d(ctx, data, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt,
              rate, ot=0.0):
    """hessian_f: the full Hessian-vector product H*dc as (3, nnz, nnx)."""
    arrs = [f32(a).ravel() for a in (data, c11, c13, c33, dc11, dc13, dc33, wavelet, taper)]
    sloc, rloc = _i32(sloc), _i32(rloc)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    gc = np.zeros((3, nnz, nnx), np.float32)
    check(_lib.vti2d_hessian(ctx.h, ptr(gc[0]), ptr(gc[1]), ptr(gc[2]), *[ptr(a) for a in arrs[:8]], ptr(sloc),
                             sloc.size // 4, ptr(rloc), rloc.size // 2, ptr(arrs[8]), nx, nz, nt, npad, dx, dz, dt,
                             rate, ot), "vti2d_hessian")
    return gc


def rtm2d(ctx, data, c11, c13, c33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate, ot=0.0,
          nhx=-1):
    """rtm_f (nhx < 0): the image (nnz, nnx); odcig_f (nhx >= 0): the offset
    gathers (2*nhx+1, nnz, nnx), offset -nhx first."""
    c11, c13, c33, wavelet, taper, data = (f32(a).ravel() for a in (c11, c13, c33, wavelet, taper, data))
    sloc, rloc = _i32(sloc), _i32(rloc)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    if nhx < 0:
        img = np.zeros((nnz, nnx), np.float32)
        check(_lib.vti2d_rtm(ctx.h, ptr(img), ptr(data), ptr(c11), ptr(c13), ptr(c33), ptr(wavelet), ptr(sloc),
                             sloc.size // 4, ptr(rloc), rloc.size // 2, ptr(taper), nx, nz, nt, npad, dx, dz, dt, rate,
                             ot), "vti2d_rtm")
    else:
        img = np.zeros((2 * nhx + 1, nnz, nnx), np.float32)
        check(_lib.vti2d_odcig(ctx.h, ptr(img), ptr(data), ptr(c11), ptr(c13), ptr(c33), ptr(wavelet), ptr(sloc),
                               sloc.size // 4, ptr(rloc), rloc.size // 2, ptr(taper), nx, nz, nhx, nt, npad, dx, dz, dt,
                               rate, ot), "vti2d_odcig")
    return img


def gn_hessian2d(ctx, c11, c13, c33, dc11, dc13, dc33, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate,
                 ot=0.0):
    """GNhessian_f: returns H*dc as (3, nnz, nnx)."""
    arrs = [f32(a).ravel() for a in (c11, c13, c33, dc11, dc13, dc33, wavelet, taper)]
    sloc, rloc = _i32(sloc), _i32(rloc)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    gc = np.zeros((3, nnz, nnx), np.float32)
    check(_lib.vti2d_gn_hessian(ctx.h, ptr(gc[0]), ptr(gc[1]), ptr(gc[2]), None, *[ptr(a) for a in arrs[:7]],
                                ptr(sloc), sloc.size // 4, ptr(rloc), rloc.size // 2, ptr(arrs[7]), nx, nz, nt, npad,
                                dx, dz, dt, rate, ot), "vti2d_gn_hessian")
    return gc


def check_eps_del(eps, dl, eps0=1.0, del0=1.0):
    """checkEpsDel on copies: returns (eps, del, mask, count)."""
    e, d = f32(eps).copy(), f32(dl).copy()
    m = np.zeros_like(e)
    n = check(_lib.vti_check_eps_del(ptr(e), ptr(d), eps0, del0, e.size, ptr(m)), "vti_check_eps_del")
    return e, d, m, n


def objfunc_gradient2d_vepsdel(ctx, data, v, eps, dl, v0, eps0, del0, wavelet, sloc, rloc, taper, nx, nz, nt, npad,
                               dx, dz, dt, rate, ot=0.0, wbottom=0.0):
    """objFuncGradientVEpsDel: returns (objective, (gv, geps, gdel), eps', del', mask); eps/del are
    projected by checkEpsDel as the reference does (the caller's arrays are not modified)."""
    v, e, d = f32(v).ravel().copy(), f32(eps).ravel().copy(), f32(dl).ravel().copy()
    w, t, dd = f32(wavelet).ravel(), f32(taper).ravel(), f32(data).ravel()
    sloc, rloc = _i32(sloc), _i32(rloc)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    g = np.zeros((3, nnz, nnx), np.float32)
    m = np.zeros(nnx * nnz, np.float32)
    f = ctypes.c_float(0.0)
    check(_lib.vti2d_objfunc_gradient_vepsdel(ctx.h, ctypes.byref(f), ptr(g[0]), ptr(g[1]), ptr(g[2]), ptr(dd), ptr(v),
                                              ptr(e), ptr(d), v0, eps0, del0, ptr(w), ptr(sloc), sloc.size // 4,
                                              ptr(rloc), rloc.size // 2, ptr(t), nx, nz, nt, npad, dx, dz, dt, rate,
                                              ot, wbottom, ptr(m)), "vti2d_objfunc_gradient_vepsdel")
    return f.value, g, e, d, m


def gn_hessian2d_vepsdel(ctx, v, eps, dl, dv, deps, ddel, v0, eps0, del0, wavelet, sloc, rloc, taper, nx, nz, nt,
                         npad, dx, dz, dt, rate, ot=0.0, wbottom=0.0):
    """GNhessianVEpsDel: returns (H (3, nnz, nnx) for (dv, deps, ddel), eps', del', mask)."""
    v, e, d = f32(v).ravel().copy(), f32(eps).ravel().copy(), f32(dl).ravel().copy()
    pv, pe, pd = (f32(a).ravel() for a in (dv, deps, ddel))
    w, t = f32(wavelet).ravel(), f32(taper).ravel()
    sloc, rloc = _i32(sloc), _i32(rloc)
    nnx, nnz = nx + 2 * npad, nz + 2 * npad
    g = np.zeros((3, nnz, nnx), np.float32)
    m = np.zeros(nnx * nnz, np.float32)
    check(_lib.vti2d_gn_hessian_vepsdel(ctx.h, ptr(g[0]), ptr(g[1]), ptr(g[2]), None, ptr(v), ptr(e), ptr(d), ptr(pv),
                                        ptr(pe), ptr(pd), v0, eps0, del0, ptr(w), ptr(sloc), sloc.size // 4, ptr(rloc),
                                        rloc.size // 2, ptr(t), nx, nz, nt, npad, dx, dz, dt, rate, ot, wbottom,
                                        ptr(m)), "vti2d_gn_hessian_vepsdel")
    return g, e, d, m


def dvepsdel_to_dcij(v, eps, dl, dv, deps, ddel, v0=1.0, eps0=1.0, del0=1.0):
    a = [f32(x).ravel() for x in (v, eps, dl, dv, deps, ddel)]
    out = [np.empty_like(a[0]) for _ in range(6)]
    check(_lib.vti_dvepsdel_to_dcij(*[ptr(o) for o in out], *[ptr(x) for x in a], v0, eps0, del0, a[0].size),
          "vti_dvepsdel_to_dcij")
    return out


def gradcij_to_gradvepsdel(gc11, gc13, gc33, v, eps, dl, v0=1.0, eps0=1.0, del0=1.0):
    a = [f32(x).ravel() for x in (gc11, gc13, gc33, v, eps, dl)]
    out = [np.empty_like(a[0]) for _ in range(3)]
    check(_lib.vti_gradcij_to_gradvepsdel(*[ptr(o) for o in out], *[ptr(x) for x in a], v0, eps0, del0, a[0].size),
          "vti_gradcij_to_gradvepsdel")
    return out


class LBFGS:
    """lbfgs (lib/lbfgs/lbfgs_c.cpp) by reverse communication, FP64 state.

    >>> opt = LBFGS(n, m=5); iflag = 0
    >>> while True:
    ...     f, g = evaluate(x)            # x: float32 array, updated in place
    ...     iflag = opt.step(x, f, g, iflag)
    ...     if iflag != 1: break
    """

    def __init__(self, n: int, m: int = 5, unit_step: bool = False, ctx=None):
        """ctx: keep the method's FP64 vectors in that context's device memory
        (vti_lbfgs_create_device); x and g stay host arrays here."""
        h = c_void_p()
        if ctx is None:
            check(_lib.vti_lbfgs_create(ctypes.byref(h), n, m, int(unit_step)), "vti_lbfgs_create")
        else:
            check(_lib.vti_lbfgs_create_device(ctypes.byref(h), ctx.h, n, m, int(unit_step)),
                  "vti_lbfgs_create_device")
        self.h = h
        self.iter = 0
        self.nfun = 0

    def step(self, x: np.ndarray, f: float, g: np.ndarray, iflag: int) -> int:
        if x.dtype != np.float32 or not x.flags["C_CONTIGUOUS"]:
            raise ValueError("x must be a contiguous float32 array (updated in place)")
        g = f32(g).ravel()
        fl, it, nf = c_int(iflag), c_int(0), c_int(0)
        check(_lib.vti_lbfgs_step(self.h, ptr(x), float(f), ptr(g), ctypes.byref(fl), ctypes.byref(it),
                                  ctypes.byref(nf)), "vti_lbfgs_step")
        self.iter, self.nfun = it.value, nf.value
        if fl.value < 0:
            raise VtiError(last_error())
        return fl.value

    def close(self):
        if self.h:
            _lib.vti_lbfgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SNAP_FORMATS = {"fp32": 0, "fp16": 1, "bf16": 2, "s16": 3}


def tune_batch2d(ctx, shots, rloc, NX, nz, nt, npad, dx, dz, dt, rate):
    """The launch shape vti2d_batch_tune picks for these shots: a dict of the
    objfunc_gradient2d_batch keyword arguments (levels, tile_rows, max_wave)
    and the winning trial's time."""
    shots, rloc = _i32(shots), _i32(rloc)
    lv, rows, wv = c_int(), c_int(), c_int()
    ms = ctypes.c_double()
    check(_lib.vti2d_batch_tune(ctx.h, ptr(shots), shots.size // 6, ptr(rloc), rloc.size // 2, NX, nz, nt, npad, dx,
                                dz, dt, rate, ctypes.byref(lv), ctypes.byref(rows), ctypes.byref(wv),
                                ctypes.byref(ms)), "vti2d_batch_tune")
    return {"levels": lv.value, "tile_rows": rows.value, "max_wave": wv.value, "trial_ms": ms.value}


def objfunc_gradient2d_batch(ctx, data, c11, c13, c33, wavelet, shots, rloc, NX, nz, nt, npad, dx, dz, dt, rate,
                             max_wave=0, levels=0, tile_rows=0, snap="fp32"):
    """Batched objective + gradient: shots = (ns, 6) ints {c0, nx, ix, iz, nr, r0};
    data = the shots' traces concatenated.  levels (1 or 2 per launch) and
    tile_rows (4, 8, 16) change only speed; snap = "fp32" (bit-exact with the
    one-shot driver), "fp16" or "bf16" (per-slice scaled 16-bit snapshots).
    Returns (objective, gc (3, nnz, NNX) float64)."""
    c11, c13, c33, wavelet, data = (f32(a).ravel() for a in (c11, c13, c33, wavelet, data))
    shots, rloc = _i32(shots), _i32(rloc)
    NNX, nnz = NX + 2 * npad, nz + 2 * npad
    gc = np.zeros((3, nnz, NNX), np.float64)
    f = ctypes.c_double(0.0)
    check(_lib.vti2d_objfunc_gradient_batch(ctx.h, ctypes.byref(f), ptr(gc[0]), ptr(gc[1]), ptr(gc[2]), ptr(data),
                                            ptr(c11), ptr(c13), ptr(c33), ptr(wavelet), ptr(shots), shots.size // 6,
                                            ptr(rloc), rloc.size // 2, NX, nz, nt, npad, dx, dz, dt, rate, max_wave,
                                            levels, tile_rows, SNAP_FORMATS[snap]),
          "vti2d_objfunc_gradient_batch")
    return f.value, gc


class BatchPlan:
    """vti2d_batch_plan_*: the batched gradient of fixed shots (geometry,
    observed data, wavelet) captured once into a hipGraph; run(c11, c13, c33)
    replays it on a new model and returns what objfunc_gradient2d_batch
    returns for the same arguments."""

    def __init__(self, ctx, data, wavelet, shots, rloc, NX, nz, nt, npad, dx, dz, dt, rate, max_wave=0, levels=0,
                 tile_rows=0, snap="fp32"):
        wavelet, data = f32(wavelet).ravel(), f32(data).ravel()
        shots, rloc = _i32(shots), _i32(rloc)
        self.shape = (3, nz + 2 * npad, NX + 2 * npad)
        h = ctypes.c_void_p()
        check(_lib.vti2d_batch_plan_create(ctx.h, ctypes.byref(h), ptr(data), ptr(wavelet), ptr(shots),
                                           shots.size // 6, ptr(rloc), rloc.size // 2, NX, nz, nt, npad, dx, dz, dt,
                                           rate, max_wave, levels, tile_rows, SNAP_FORMATS[snap]),
              "vti2d_batch_plan_create")
        self.h, self._ctx = h, ctx

    def run(self, c11, c13, c33):
        c11, c13, c33 = (f32(a).ravel() for a in (c11, c13, c33))
        gc = np.zeros(self.shape, np.float64)
        f = ctypes.c_double(0.0)
        check(_lib.vti2d_batch_plan_run(self.h, ctypes.byref(f), ptr(gc[0]), ptr(gc[1]), ptr(gc[2]), ptr(c11),
                                        ptr(c13), ptr(c33)), "vti2d_batch_plan_run")
        return f.value, gc

    def close(self):
        if self.h:
            check(_lib.vti2d_batch_plan_destroy(self.h), "vti2d_batch_plan_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Ticket:
    """A submitted server request: its id, its output arrays and the inputs
    the server reads until it is done."""

    def __init__(self, rid, kind, out, keep):
        self.id, self.kind, self.out, self._keep = rid, kind, out, keep


class Server:
    """Request server on one device (vti_serve_*): `slots` concurrent
    workers, admission by device-memory budget, resident models, 2D
    gradient requests of one schedule merged into batched runs (at most
    `batch` per run; 1 = off).  Submit with grad2d / model3d, collect with
    wait."""

    def __init__(self, device=0, slots=4, budget_bytes=0.0, batch=8, snap="fp32", exclusive3d=True):
        h = ctypes.c_void_p()
        check(_lib.vti_serve_create(ctypes.byref(h), device, slots, float(budget_bytes)), "vti_serve_create")
        self.h = h
        self._models = {}
        # the host arrays of every request not yet waited for: the server reads
        # its inputs and writes its outputs until the request is done, so they
        # stay alive here even if the caller drops the Ticket
        self._pending = {}
        check(_lib.vti_serve_set_batching(self.h, batch), "vti_serve_set_batching")
        check(_lib.vti_serve_set_snapshots(self.h, SNAP_FORMATS[snap]), "vti_serve_set_snapshots")
        check(_lib.vti_serve_set_exclusive3d(self.h, int(bool(exclusive3d))), "vti_serve_set_exclusive3d")

    def close(self):
        if self.h:
            check(_lib.vti_serve_destroy(self.h), "vti_serve_destroy")  # drains the queue
            self.h = None
            self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def put_model(self, c11, c13, c33):
        """Upload c11, c13, c33 once; returns the model handle."""
        c = [f32(a).ravel() for a in (c11, c13, c33)]
        assert c[0].size == c[1].size == c[2].size
        hd = c_int()
        check(_lib.vti_serve_put_model(self.h, ptr(c[0]), ptr(c[1]), ptr(c[2]), c[0].size, ctypes.byref(hd)),
              "vti_serve_put_model")
        self._models[hd.value] = c[0].size
        return hd.value

    def drop_model(self, handle):
        check(_lib.vti_serve_drop_model(self.h, handle), "vti_serve_drop_model")
        self._models.pop(handle, None)

    def grad2d(self, model, data, wavelet, sloc, rloc, taper, nx, nz, nt, npad, dx, dz, dt, rate):
        """objFuncGradient_f on a resident model; wait() returns (objective, gc (3, nnz, nnx))."""
        data, wavelet, taper = (f32(a).ravel() for a in (data, wavelet, taper))
        sloc, rloc = _i32(sloc), _i32(rloc)
        gc = np.zeros((3, nz + 2 * npad, nx + 2 * npad), np.float32)
        f = np.zeros(1, np.float32)
        rid = c_longlong()
        check(_lib.vti_serve_grad2d(self.h, ctypes.byref(rid), model, ptr(f), ptr(gc[0]), ptr(gc[1]), ptr(gc[2]),
                                    ptr(data), ptr(wavelet), ptr(sloc), sloc.size // 4, ptr(rloc), rloc.size // 2,
                                    ptr(taper), nx, nz, nt, npad, dx, dz, dt, rate), "vti_serve_grad2d")
        self._pending[rid.value] = ((f, gc), (data, wavelet, taper, sloc, rloc))
        return Ticket(rid.value, "grad2d", (f, gc), (data, wavelet, taper, sloc, rloc))

    def model3d(self, model, souloc, shotid, recloc, wavelet, shape, nt, npad, o, d, dt, rate):
        """modelDataCij3d_f on a resident model of shape (nz, ny, nx); wait()
        returns the trace-major data (nr_total * nnt)."""
        souloc, shotid, recloc, wavelet = _geom3(souloc, shotid, recloc, wavelet)
        nz, ny, nx = shape
        sts = int(np.round(np.float32(rate) / np.float32(dt)))
        nnt = (nt - 1) // sts + 1
        nrt = recloc.size // 4
        dd = np.zeros(nrt * nnt, np.float32)
        rid = c_longlong()
        check(_lib.vti_serve_model3d(self.h, ctypes.byref(rid), model, ptr(dd), ptr(souloc), souloc.size // 5,
                                     ptr(shotid), shotid.size, ptr(recloc), nrt, ptr(wavelet), nx, ny, nz, nt, npad,
                                     *o, 0.0, *d, dt, rate), "vti_serve_model3d")
        self._pending[rid.value] = (dd, (souloc, shotid, recloc, wavelet))
        return Ticket(rid.value, "model3d", dd, (souloc, shotid, recloc, wavelet))

    def wait(self, ticket):
        """Block until the request is done: (result, latency_s, service_s)."""
        lat, svc = ctypes.c_double(), ctypes.c_double()
        rc = _lib.vti_serve_wait(self.h, ticket.id, ctypes.byref(lat), ctypes.byref(svc))
        self._pending.pop(ticket.id, None)  # the request is done, whatever its status
        check(rc, "vti_serve_wait")
        if ticket.kind == "grad2d":
            f, gc = ticket.out
            res = (float(f[0]), gc)
        else:
            res = ticket.out
        ticket._keep = None
        return res, lat.value, svc.value

    def stats(self):
        p50, p99, rate, peak = (ctypes.c_double() for _ in range(4))
        n = c_longlong()
        check(_lib.vti_serve_stats(self.h, ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(rate), ctypes.byref(n),
                                   ctypes.byref(peak)), "vti_serve_stats")
        runs, merged = c_longlong(), c_longlong()
        check(_lib.vti_serve_batch_stats(self.h, ctypes.byref(runs), ctypes.byref(merged)), "vti_serve_batch_stats")
        return {"p50_s": p50.value, "p99_s": p99.value, "req_per_s": rate.value, "completed": n.value,
                "peak_bytes": peak.value, "merged_runs": runs.value, "merged_requests": merged.value}
