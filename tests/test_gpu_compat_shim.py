"""The drop-in boundary executed: the reference's C++ driver signatures
(include/vti_compat.hpp, compiled into tests/libcompat_harness.so) called the
way the reference's mains call them, against direct calls of the C ABI.

- synthetic_f, objFuncGradient_f, objFuncGradientCij3d_f, rtmCij3d_f:
  bit-equal to vti2d_synthetic, vti2d_objfunc_gradient,
  vti3d_objfunc_gradient, vti3d_rtm (the shim's argument marshalling: shot
  and GPU vectors, the observed data, cij's shell left holding the random
  boundary);
- the group forms objFuncGradientCij3d / objFuncCij3d / modelData3d_f
  (wave3d.h:94,98,108): the reference's shot split and FP32 sums over groups;
- the 3D objFuncGradientVEpsDel and rtmVEpsDel (objFuncGradient3d_f.cu:249,
  rtm3d_f.cu:139): the composition of the pinned host helpers and the C ABI.
"""
import ctypes
import os

import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel
from common3d_drivers import problem

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def H(vti):
    path = os.path.join(HERE, "libcompat_harness.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -C thesis-source-code_amd)")
    h = ctypes.CDLL(path)
    h.h_objFuncGradient_f.restype = ctypes.c_float
    h.h_objFuncGradientVEpsDel3d.restype = ctypes.c_double
    return h


def P_(a):
    return a.ctypes.data_as(ctypes.c_void_p)


F = ctypes.c_float


def _boundaries(vti, P):
    nx, ny, nz, npad = P["nx"], P["ny"], P["nz"], P["npad"]
    nb = P["nbnd"]
    padb = np.zeros(3 * nb, np.float32)
    for k in range(3):
        vti.check(vti.lib.vti_get_boundary(vti.ptr(padb[k * nb:(k + 1) * nb]), vti.ptr(P["cij"][k]), nx, ny, nz, npad))
    randb = (padb * np.tile(P["shell_scale"], 3)).astype(np.float32)
    return padb, randb


def _obs(vti, ctx, P):
    return vti.model_data3d(ctx, P["souloc"], np.arange(len(P["souloc"])), P["recloc"], P["w"], *P["ctrue"], P["nt"],
                            P["npad"], P["o"], P["d"], P["dt"], P["rate"])


def test_shim_2d_bitexact(vti, ctx, H):
    P = problem2d(vti, ns=2)
    c = [np.ascontiguousarray(a, np.float32) for a in P["c"]]
    w, sloc, rloc, taper = P["w"], P["sloc"], P["rloc"], np.ascontiguousarray(P["taper"], np.float32)
    print("shim 2D: synthetic_f, objFuncGradient_f", flush=True)
    d_abi = vti.synthetic2d(ctx, *c, w, sloc, rloc, taper, *args2d(P))
    d_shim = np.zeros_like(d_abi)
    a = args2d(P)
    H.h_synthetic_f(P_(d_shim), *[P_(x) for x in c], P_(w), P_(sloc), 2, P_(rloc), len(rloc), P_(taper), a[0], a[1],
                    a[2], a[3], F(a[4]), F(a[5]), F(a[6]), F(a[7]))
    assert np.array_equal(d_shim, d_abi)
    obs = (d_abi * np.float32(0.9)).astype(np.float32)
    f_abi, g_abi = vti.objfunc_gradient2d(ctx, obs, *c, w, sloc, rloc, taper, *a)
    g = np.zeros_like(g_abi)
    f = H.h_objFuncGradient_f(P_(g[0]), P_(g[1]), P_(g[2]), P_(obs), *[P_(x) for x in c], P_(w), P_(sloc), 2,
                              P_(rloc), len(rloc), P_(taper), a[0], a[1], a[2], a[3], F(a[4]), F(a[5]), F(a[6]),
                              F(a[7]))
    assert f == np.float32(f_abi)
    assert np.array_equal(g, g_abi)
    print("shim 2D: modeling_f", flush=True)
    sxz = int(sloc[0, 0] + sloc[0, 1] * P["nnx"])
    m_abi = vti.modeling2d(ctx, *c, w, sxz, taper, *a)
    m_shim = np.zeros_like(m_abi)
    H.h_modeling_f(P_(m_shim), *[P_(x) for x in c], P_(w), sxz, P_(taper), a[0], a[1], a[2], a[3], F(a[4]), F(a[5]),
                   F(a[6]), F(a[7]))
    assert np.abs(m_abi).max() > 0
    assert np.array_equal(m_shim, m_abi)


def _g3(P):
    return (P["nx"], P["ny"], P["nz"], P["nt"], P["npad"], F(P["o"][0]), F(P["d"][0]), F(P["dt"]), F(P["rate"]))


def test_shim_3d_gradient_and_rtm_bitexact(vti, ctx, H):
    P = problem()
    obs = _obs(vti, ctx, P)
    padb, randb = _boundaries(vti, P)
    sou, rec, w = (np.ascontiguousarray(P[k], np.float32) for k in ("souloc", "recloc", "w"))
    shots = np.array([1], np.int32)
    print("shim 3D: objFuncGradientCij3d_f", flush=True)
    c_abi = P["cij"].copy()
    f_abi, g_abi = vti.objfunc_gradient3d(ctx, sou, shots, rec, w, c_abi, padb, randb, P["nt"], P["npad"], P["o"],
                                          P["d"], P["dt"], P["rate"], obs)
    c_shim = P["cij"].copy()
    fg = np.zeros(1 + c_shim.size, np.float32)
    H.h_objFuncGradientCij3d_f(P_(fg), P_(sou), len(sou), P_(shots), 1, P_(rec), P_(w), P_(c_shim), P_(padb),
                               P_(randb), *_g3(P), 1, P_(obs))
    assert fg[0] == np.float32(f_abi)
    assert np.array_equal(fg[1:].reshape(g_abi.shape), g_abi)
    assert np.array_equal(c_shim, c_abi)  # the shell left holding the random boundary
    print("shim 3D: rtmCij3d_f", flush=True)
    rec3 = np.ascontiguousarray(rec[:, :3])
    img_abi = vti.rtm3d(ctx, sou, shots, rec3, w, *P["cij"], P["nt"], P["npad"], P["o"], P["d"], P["dt"], P["rate"],
                        obs)
    img = np.zeros_like(img_abi)
    H.h_rtmCij3d_f(P_(img), P_(sou), len(sou), P_(shots), 1, P_(rec3), P_(w), *[P_(P["cij"][k]) for k in range(3)],
                   *_g3(P), 1, P_(obs))
    assert np.abs(img_abi).max() > 0
    assert np.array_equal(img, img_abi)


def test_shim_3d_groups(vti, ctx, H):
    """ngpugroup = 2 over shots {0, 1} and GPUs {0, 1}: group 0 runs shot 0 on
    GPU 0, group 1 shot 1 on GPU 1, each on its own copy of cij; fgcij = 0 +
    group 1 + group 0 in FP32 (objFuncGradient3d_f.cu:307-359); objFuncCij3d
    adds the groups' objectives to *f in group order."""
    P = problem()
    obs = _obs(vti, ctx, P)
    padb, randb = _boundaries(vti, P)
    sou, rec, w = (np.ascontiguousarray(P[k], np.float32) for k in ("souloc", "recloc", "w"))
    per = []
    for s in (0, 1):
        per.append(vti.objfunc_gradient3d(ctx, sou, np.array([s]), rec, w, P["cij"].copy(), padb, randb, P["nt"],
                                          P["npad"], P["o"], P["d"], P["dt"], P["rate"], obs))
    want = np.zeros(1 + P["cij"].size, np.float32)
    for s in (1, 0):
        want += np.concatenate([[np.float32(per[s][0])], per[s][1].ravel()]).astype(np.float32)
    cij = P["cij"].copy()
    fg = np.zeros_like(want)
    shots = np.array([0, 1], np.int32)
    print("shim 3D: objFuncGradientCij3d, 2 groups", flush=True)
    H.h_objFuncGradientCij3d(P_(fg), P_(sou), len(sou), P_(shots), 2, P_(rec), P_(w), P_(cij), P_(padb), P_(randb),
                             *_g3(P), 2, 2, P_(obs))
    assert np.array_equal(fg, want)
    assert np.array_equal(cij, P["cij"])  # groups work on copies
    f = np.array([0.25], np.float32)
    fo = [vti.objfunc_gradient3d(ctx, sou, np.array([s]), rec, w, P["cij"].copy(), padb, None, P["nt"], P["npad"],
                                 P["o"], P["d"], P["dt"], P["rate"], obs, gradient=False) for s in (0, 1)]
    cij2 = P["cij"].copy()
    H.h_objFuncCij3d(P_(f), P_(sou), len(sou), P_(shots), 2, P_(rec), P_(w), P_(cij2), P_(padb), *_g3(P), 2, 2,
                     P_(obs))
    assert f[0] == (np.float32(0.25) + np.float32(fo[0])) + np.float32(fo[1])
    print("shim 3D: modelData3d_f, 2 groups", flush=True)
    dd_abi = vti.model_data3d(ctx, sou, shots, rec, w, *P["ctrue"], P["nt"], P["npad"], P["o"], P["d"], P["dt"],
                              P["rate"])
    dd = np.zeros_like(dd_abi)
    H.h_modelData3d_f(P_(dd), P_(sou), len(sou), P_(shots), 2, P_(rec), P_(w), *[P_(P["ctrue"][k]) for k in range(3)],
                      *_g3(P), 2, 2)
    assert np.array_equal(dd, dd_abi)


def test_shim_3d_vepsdel(vti, ctx, H):
    """objFuncGradientVEpsDel (3D) and rtmVEpsDel equal their compositions of
    the pinned host helpers and the C-ABI drivers, bit for bit."""
    from common3d import layered_model

    P = problem()
    nx, ny, nz, npad = P["nx"], P["ny"], P["nz"], P["npad"]
    v, e, dl = layered_model(nz, ny, nx, seed=11)
    v0, eps0 = 1000.0, 0.2
    vs, es = (v / np.float32(v0)).astype(np.float32), (e / np.float32(eps0)).astype(np.float32)
    dls = dl.astype(np.float32)
    nb = P["nbnd"]

    def shell(a):
        b = np.zeros(nb, np.float32)
        vti.check(vti.lib.vti_get_boundary(vti.ptr(b), vti.ptr(np.ascontiguousarray(a)), nx, ny, nz, npad))
        return b

    padv = shell(vs)
    randv = (padv * P["shell_scale"]).astype(np.float32)
    beps, bdel = shell(es), shell(dls)
    c = vti.vepsdel_to_cij(vs, es, dls, v0, eps0, 1.0)
    padc = vti.vepsdel_to_cij(padv, beps, bdel, v0, eps0, 1.0)
    randc = vti.vepsdel_to_cij(randv, beps, bdel, v0, eps0, 1.0)
    obs = _obs(vti, ctx, P)
    sou, rec, w = (np.ascontiguousarray(P[k], np.float32) for k in ("souloc", "recloc", "w"))
    shots = np.array([0], np.int32)
    cij = np.stack([np.asarray(x, np.float32).reshape(nz, ny, nx) for x in c]).astype(np.float32)
    f_abi, gc = vti.objfunc_gradient3d(ctx, sou, shots, rec, w, cij, np.concatenate(padc), np.concatenate(randc),
                                       P["nt"], npad, P["o"], P["d"], P["dt"], P["rate"], obs)
    wbottom = 10 * P["d"][2]
    nw = int(np.float32(np.float32(wbottom) - np.float32(0.0)) / np.float32(P["d"][2]) + 1) - npad
    gc = gc.copy()
    for k in range(3):
        vti.check(vti.lib.vti_zero_boundary(vti.ptr(gc[k]), nx, ny, nz, npad))
        gc[k][npad:npad + nw] = 0
    want = vti.gradcij_to_gradvepsdel(gc[0], gc[1], gc[2], vs, es, dls, v0, eps0, 1.0)
    gv, ge, gd = (np.zeros((nz, ny, nx), np.float32) for _ in range(3))
    print("shim 3D: objFuncGradientVEpsDel", flush=True)
    vc, ec, dc = vs.copy(), es.copy(), dls.copy()
    f = H.h_objFuncGradientVEpsDel3d(P_(gv), P_(ge), P_(gd), P_(sou), len(sou), P_(shots), 1, P_(rec), P_(w), P_(vc),
                                     P_(ec), P_(dc), P_(padv), P_(randv), *_g3(P), F(v0), F(eps0), F(wbottom), 1, 1,
                                     P_(obs))
    assert f == float(np.float32(f_abi))
    for got, exp in zip((gv, ge, gd), want):
        assert np.array_equal(got.ravel(), np.asarray(exp, np.float32).ravel())
    print("shim 3D: rtmVEpsDel", flush=True)
    c1 = vti.vepsdel_to_cij(vs, es, dls, 1.0, 1.0, 1.0)
    rec3 = np.ascontiguousarray(rec[:, :3])
    img_abi = vti.rtm3d(ctx, sou, shots, rec3, w, *[np.asarray(x, np.float32).reshape(nz, ny, nx) for x in c1],
                        P["nt"], npad, P["o"], P["d"], P["dt"], P["rate"], obs)
    vti.check(vti.lib.vti_zero_boundary(vti.ptr(img_abi), nx, ny, nz, npad))
    img_abi[npad:npad + nw] = 0
    img = np.zeros_like(img_abi)
    vc, ec, dc = vs.copy(), es.copy(), dls.copy()
    H.h_rtmVEpsDel(P_(img), P_(sou), len(sou), P_(shots), 1, P_(rec3), P_(w), P_(vc), P_(ec), P_(dc), *_g3(P),
                   F(wbottom), 1, P_(obs))
    assert np.abs(img).max() > 0
    assert np.array_equal(img, img_abi)
    assert rel(img, img_abi) == 0.0
