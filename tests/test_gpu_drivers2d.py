"""2D drivers on the GPU against the oracle (lib/vti-fwi-2d restated).

synthetic_f, objFuncGradient_f and GNhessian_f are bit-exact with the
oracle's FP32 restatement of the reference (one shot per GPU, so that the
cross-shot sum is one rounding as in the HIP path's FP64 sum of two shots),
and each result states its error against the FP64 oracle."""
import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def P(vti):
    return problem2d(vti)


@pytest.fixture(scope="module")
def obs(orc, P):
    return orc.synthetic2d(orc.F64, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)).astype(
        np.float32)


def test_synthetic_bitexact(vti, orc, ctx, P):
    _say("2d synthetic", P["nnz"], P["nnx"], P["nt"])
    d = vti.synthetic2d(ctx, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    o = orc.synthetic2d(orc.F32, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    o64 = orc.synthetic2d(orc.F64, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    assert np.abs(o).max() > 0
    assert np.array_equal(d, o)
    assert rel(d, o64) < 4e-5  # the reference's FP32 itself: 2.0e-5


def test_synthetic_ratio1_and_negative_time(vti, orc, ctx, P):
    """rate == dt records level 1 too; ot < 0 shifts the recorded window."""
    for rate, ot in ((1e-3, 0.0), (4e-3, -0.02)):
        _say("2d synthetic rate", rate, "ot", ot)
        a = list(args2d(P))
        a[-1] = rate
        d = vti.synthetic2d(ctx, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *a, ot=ot)
        o = orc.synthetic2d(orc.F32, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *a, ot=ot)
        assert np.abs(o).max() > 0
        assert np.array_equal(d, o)


@pytest.mark.parametrize("rate,ot", [(4e-3, 0.0), (1e-3, 0.0), (4e-3, -0.02)])
def test_modeling_bitexact(vti, orc, ctx, P, rate, ot):
    """modeling_f: the p field of every recorded slot equals the oracle's FP32
    restatement bit for bit, slots never recorded keep their input values,
    and the FP64 oracle states the error."""
    _say("2d modeling rate", rate, "ot", ot)
    a = list(args2d(P))
    a[-1] = rate
    sxz = int(P["sloc"][0, 0] + P["sloc"][0, 1] * P["nnx"])
    o = orc.modeling2d(orc.F32, *P["c"], P["w"], sxz, P["taper"], *a, ot=ot)
    init = np.full(o.shape, 3.0, np.float32)
    o = orc.modeling2d(orc.F32, *P["c"], P["w"], sxz, P["taper"], *a, ot=ot, init=init)
    m = vti.modeling2d(ctx, *P["c"], P["w"], sxz, P["taper"], *a, ot=ot, wavefield=init.copy())
    o64 = orc.modeling2d(orc.F64, *P["c"], P["w"], sxz, P["taper"], *a, ot=ot, init=init)
    assert np.abs(o[o != 3.0]).max() > 0
    assert np.array_equal(m, o)
    e = rel(m[m != 3.0], o64[m != 3.0].astype(np.float32))
    _say(f"  rel vs FP64 {e:.3e}")
    assert e < 1e-4


@pytest.mark.parametrize("nshot", [1, 2])
@pytest.mark.parametrize("kind", ["zero", "model"])
def test_objfunc_gradient(vti, orc, ctx, P, obs, nshot, kind):
    _say("2d gradient nshot", nshot, kind)
    sloc = P["sloc"][:nshot]
    data = obs if kind == "model" else np.zeros_like(obs)
    f, g = vti.objfunc_gradient2d(ctx, data, *P["c"], P["w"], sloc, P["rloc"], P["taper"], *args2d(P))
    fo, go = orc.fg2d(orc.F32, False, data, *P["c"], None, P["w"], sloc, P["rloc"], P["taper"], *args2d(P),
                      ngpu=nshot)
    f64, g64 = orc.fg2d(orc.F64, False, data, *P["c"], None, P["w"], sloc, P["rloc"], P["taper"], *args2d(P))
    assert np.abs(go).max() > 0
    assert np.array_equal(g, go)
    assert np.float32(f) == np.float32(fo)
    e = rel(g, g64)
    _say("  rel vs FP64", e)
    # zero data: well conditioned (1.1e-5).  Model data: the residual is 0.2% of
    # the data, so FP32 trace rounding shows at 1e-2 of it (as in the reference)
    assert e < (1e-4 if kind == "zero" else 3e-2)


@pytest.mark.parametrize("nshot", [1, 2])
def test_gn_hessian(vti, orc, ctx, P, nshot):
    _say("2d GN Hv nshot", nshot)
    rng = np.random.default_rng(7)
    dc = [(rng.standard_normal((P["nnz"], P["nnx"])) * P["c"][k].mean() * 0.01).astype(np.float32) for k in range(3)]
    sloc = P["sloc"][:nshot]
    h = vti.gn_hessian2d(ctx, *P["c"], *dc, P["w"], sloc, P["rloc"], P["taper"], *args2d(P))
    ho = orc.fg2d(orc.F32, True, None, *P["c"], dc, P["w"], sloc, P["rloc"], P["taper"], *args2d(P), ngpu=nshot)[1]
    h64 = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], sloc, P["rloc"], P["taper"], *args2d(P), ngpu=nshot)[1]
    assert np.abs(ho).max() > 0
    assert np.array_equal(h, ho)
    e = rel(h, h64)
    _say("  rel vs FP64", e)
    assert e < 1e-4


def test_rejects_unaligned_schedule(vti, ctx, P, obs):
    a = list(args2d(P))
    a[2] = P["nt"] + 1  # (nt-1) % ratio != 0
    with pytest.raises(vti.VtiError, match="snapshot"):
        vti.objfunc_gradient2d(ctx, obs, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *a)


def test_synthetic_cfg1_size(vti, orc, ctx):
    """256 x 256 interior + 40-point boundary (336 x 336), 2000 steps, one
    shot, 256 receivers at z=10: bit-exact with the oracle."""
    P = problem2d(vti, nx=256, nz=256, npad=40, nt=2001, ns=1, rate=4e-3, rec_step=1, seed=9)
    P["sloc"] = np.array([[40 + 128, 40 + 20, 256, 0]], np.int32)  # source at (z=20, x=128)
    P["rloc"] = np.array([[40 + ix, 40 + 10] for ix in range(256)], np.int32)  # receivers at z=10
    _say("2d synthetic 336x336x2000")
    d = vti.synthetic2d(ctx, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    o = orc.synthetic2d(orc.F32, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    assert np.abs(o).max() > 0
    assert np.array_equal(d, o)


def test_modeling_cfg1_size(vti, orc, ctx):
    """modeling_f at the 336 x 336, 2000-step size, source at (z=20, x=128),
    snapshots every 4 ms: every slot bit-exact with the oracle."""
    P = problem2d(vti, nx=256, nz=256, npad=40, nt=2001, ns=1, rate=4e-3, rec_step=1, seed=9)
    sxz = (40 + 128) + (40 + 20) * P["nnx"]
    _say("2d modeling 336x336x2000")
    m = vti.modeling2d(ctx, *P["c"], P["w"], sxz, P["taper"], *args2d(P))
    o = orc.modeling2d(orc.F32, *P["c"], P["w"], sxz, P["taper"], *args2d(P))
    assert m.shape == (501, 336, 336)
    assert np.abs(o).max() > 0
    assert np.array_equal(m, o)


def _vepsdel_problem(vti):
    from common2d import layered2d

    P = problem2d(vti)
    v, e, dl = layered2d(P["nnz"], P["nnx"], seed=3)
    v0, eps0, del0 = 2000.0, 0.5, 0.25
    vn, en, dn = v / v0, e / eps0, dl / del0
    dn = dn.copy()
    dn[P["npad"] + 20: P["npad"] + 24, :] = en[P["npad"] + 20: P["npad"] + 24, :] * eps0 / del0 * 1.3  # del > eps
    return P, vn.astype(np.float32), en.astype(np.float32), dn.astype(np.float32), (v0, eps0, del0)


def test_vepsdel_gradient_wrapper(vti, ctx):
    """objFuncGradientVEpsDel = checkEpsDel, VEpsDel2Cij, objFuncGradient_f,
    GradCij2GradVEpsDel and zeroed rows (each pinned on its own): the wrapper
    equals that composition bit for bit."""
    P, v, e, d, (v0, eps0, del0) = _vepsdel_problem(vti)
    a = args2d(P)
    wb = 30.0
    ct = vti.vepsdel_to_cij(v * 1.03, e, np.minimum(d, e * eps0 / del0), v0, eps0, del0)
    data = vti.synthetic2d(ctx, *ct, P["w"], P["sloc"], P["rloc"], P["taper"], *a)
    f, g, e2, d2, m = vti.objfunc_gradient2d_vepsdel(ctx, data, v, e, d, v0, eps0, del0, P["w"], P["sloc"], P["rloc"],
                                                      P["taper"], *a, wbottom=wb)
    ce, cd, cm, count = vti.check_eps_del(e, d, eps0, del0)
    assert count > 0
    c = vti.vepsdel_to_cij(v, ce, cd, v0, eps0, del0)
    f1, gc = vti.objfunc_gradient2d(ctx, data, *c, P["w"], P["sloc"], P["rloc"], P["taper"], *a)
    gg = vti.gradcij_to_gradvepsdel(gc[0], gc[1], gc[2], v, ce, cd, v0, eps0, del0)
    rows = P["npad"] + int(wb / P["dz"] + 1)
    for k in range(3):
        gk = gg[k].reshape(P["nnz"], P["nnx"])
        gk[:rows] = 0
        assert np.array_equal(g[k], gk)
    assert f == f1
    assert np.array_equal(e2, ce.ravel()) and np.array_equal(d2, cd.ravel()) and np.array_equal(m, cm.ravel())
    assert np.abs(g[0]).max() > 0


def test_vepsdel_gn_wrapper(vti, ctx):
    P, v, e, d, (v0, eps0, del0) = _vepsdel_problem(vti)
    a = args2d(P)
    wb = 20.0
    rng = np.random.default_rng(4)
    dv, de, dd = (rng.standard_normal(v.shape).astype(np.float32) * 0.01 for _ in range(3))
    h, e2, d2, m = vti.gn_hessian2d_vepsdel(ctx, v, e, d, dv, de, dd, v0, eps0, del0, P["w"], P["sloc"], P["rloc"],
                                             P["taper"], *a, wbottom=wb)
    ce, cd, cm, _ = vti.check_eps_del(e, d, eps0, del0)
    cc = vti.dvepsdel_to_dcij(v, ce, cd, dv, de, dd, v0, eps0, del0)
    hc = vti.gn_hessian2d(ctx, *cc, P["w"], P["sloc"], P["rloc"], P["taper"], *a)
    gg = vti.gradcij_to_gradvepsdel(hc[0], hc[1], hc[2], v, ce, cd, v0, eps0, del0)
    rows = int(np.float32(P["npad"]) + np.float32(wb) / np.float32(P["dz"]) + np.float32(1))
    for k in range(3):
        gk = gg[k].reshape(P["nnz"], P["nnx"])
        gk[:rows] = 0
        assert np.array_equal(h[k], gk)
    assert np.abs(h[0]).max() > 0


@pytest.mark.parametrize("dx,dz", [(10.0, 12.5), (1500.0, 1500.0), (2000.0, 10.0)])
def test_spacings_bitexact(vti, orc, ctx, dx, dz):
    """The 2D kernels divide by dx^2 and dz^2 through a reciprocal proven for
    each divisor, or by the division when no proof is made (above 2^20:
    1500^2 and 2000^2).  Unequal spacings, both routes and a mix of the two
    stay bit-exact with the oracle's FP32 restatement: one-shot gradients,
    the batched multi-shot route and batched modeling."""
    import ctypes

    vti.lib.vti_div_reciprocal.restype = ctypes.c_float
    vti.lib.vti_div_reciprocal.argtypes = [ctypes.c_float]
    for d in (dx, dz):
        y = vti.lib.vti_div_reciprocal(np.float32(d) * np.float32(d))
        assert (y != 0) == (d * d <= 2.0 ** 20)
    P2 = problem2d(vti, ns=2, seed=8, dx=dx)
    P2["dz"] = dz
    a = args2d(P2)
    _say("2d spacings", dx, dz)
    data = orc.synthetic2d(orc.F64, *P2["ctrue"], P2["w"], P2["sloc"], P2["rloc"], P2["taper"], *a).astype(
        np.float32) * np.float32(0.5)
    for ns in (1, 2):
        s = P2["sloc"][:ns]
        f, g = vti.objfunc_gradient2d(ctx, data, *P2["c"], P2["w"], s, P2["rloc"], P2["taper"], *a)
        fo, go = orc.fg2d(orc.F32, False, data, *P2["c"], None, P2["w"], s, P2["rloc"], P2["taper"], *a, ngpu=ns)
        assert np.abs(go).max() > 0
        assert np.array_equal(g, go)
        assert np.float32(f) == np.float32(fo)
    d = vti.synthetic2d(ctx, *P2["c"], P2["w"], P2["sloc"], P2["rloc"], P2["taper"], *a)
    o = orc.synthetic2d(orc.F32, *P2["c"], P2["w"], P2["sloc"], P2["rloc"], P2["taper"], *a)
    assert np.abs(o).max() > 0
    assert np.array_equal(d, o)


def test_multishot_call_runs_as_batch_bit_exact(vti, ctx):
    """A call with several shots and the standard taper runs them as one
    batch: its gradient equals the FP64 sum of one-shot calls rounded once,
    and with a taper that differs from the standard one only at a corner
    point no step reads (so the call takes the one-shot loop) the result is
    the same bit for bit."""
    P3 = problem2d(vti, ns=3, seed=4)
    a = args2d(P3)
    data = vti.synthetic2d(ctx, *P3["ctrue"], P3["w"], P3["sloc"], P3["rloc"], P3["taper"], *a)
    data = (data * np.float32(0.9)).astype(np.float32)
    f, g = vti.objfunc_gradient2d(ctx, data, *P3["c"], P3["w"], P3["sloc"], P3["rloc"], P3["taper"], *a)
    acc = np.zeros(g.shape, np.float64)
    fs = 0.0
    for k in range(3):
        fk, gk = vti.objfunc_gradient2d(ctx, data, *P3["c"], P3["w"], P3["sloc"][k:k + 1], P3["rloc"], P3["taper"],
                                        *a)
        acc += gk.astype(np.float64)
        fs += fk
    assert np.abs(g).max() > 0
    assert np.array_equal(g, acc.astype(np.float32))
    assert abs(f - fs) <= 1e-6 * abs(fs)
    t2 = P3["taper"].copy()
    t2.flat[0] = np.float32(0.5)
    f2, g2 = vti.objfunc_gradient2d(ctx, data, *P3["c"], P3["w"], P3["sloc"], P3["rloc"], t2, *a)
    assert np.array_equal(g2, g) and np.float32(f2) == np.float32(f)


def test_multishot_synthetic_runs_as_batch_bit_exact(vti, ctx):
    """synthetic_f with several shots and the standard taper runs them as one
    batch: the traces equal one-shot calls' bit for bit, the row the
    reference never records keeps the caller's values, and a taper that
    differs only at a corner point nothing reads (the one-shot loop) gives
    the same bits."""
    P3 = problem2d(vti, ns=3, seed=6)
    a = args2d(P3)
    nnt = (P3["nt"] - 1) // 4 + 1
    fill = np.full((nnt, P3["nr"]), -7.0, np.float32)
    d = vti.synthetic2d(ctx, *P3["ctrue"], P3["w"], P3["sloc"], P3["rloc"], P3["taper"], *a, data=fill.copy())
    ref = fill.copy()
    for k in range(3):
        dk = vti.synthetic2d(ctx, *P3["ctrue"], P3["w"], P3["sloc"][k:k + 1], P3["rloc"], P3["taper"], *a,
                             data=fill.copy())
        c0, n = P3["sloc"][k, 3], P3["sloc"][k, 2]
        ref[:, c0:c0 + n] = dk[:, c0:c0 + n]
    assert np.abs(d[1:]).max() > 0
    assert np.array_equal(d, ref)
    assert np.all(d[0] == -7.0)
    t2 = P3["taper"].copy()
    t2.flat[0] = np.float32(0.5)
    d2 = vti.synthetic2d(ctx, *P3["ctrue"], P3["w"], P3["sloc"], P3["rloc"], t2, *a, data=fill.copy())
    assert np.array_equal(d2, d)


def test_multishot_call_waves_in_flight_match_batched_driver(vti, ctx):
    """A 20-shot call runs as waves of 4 shots, two in flight on two streams
    (the forward pass of one beside the adjoint of the other).  Its gradient
    and objective equal the batched driver's on the same shots, rounded to
    FP32, and those do not depend on how the batched driver splits the
    shots: one wave of all 20, or 20 waves of one, two in flight, give the
    same FP64 bits."""
    P = problem2d(vti, ns=20, seed=11)
    a = args2d(P)
    data = vti.synthetic2d(ctx, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *a)
    data = (data * np.float32(0.9)).astype(np.float32)
    f, g = vti.objfunc_gradient2d(ctx, data, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *a)
    shots = np.array([(0, P["nx"], ix, iz, nr, r0) for ix, iz, nr, r0 in P["sloc"]], np.int32)
    cat = np.concatenate([data[:, r0:r0 + nr].ravel() for _, _, nr, r0 in P["sloc"]])
    out = [vti.objfunc_gradient2d_batch(ctx, cat, *P["c"], P["w"], shots, P["rloc"], *a, max_wave=m)
           for m in (0, 1)]
    (f0, g0), (f1, g1) = out
    assert np.abs(g0).max() > 0
    assert np.array_equal(g0, g1) and f0 == f1
    assert np.array_equal(g, g0.astype(np.float32))
    assert np.float32(f) == np.float32(f0)
