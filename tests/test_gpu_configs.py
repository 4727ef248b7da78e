"""Parity at the sizes of the benchmark configurations, where the launch
shapes, tile counts, XCD remapping, row pitch and 32-bit chunk offsets of the
real workloads are exercised (the other GPU tests use small grids).

- cfg2 (2D, 512 x 1024 + 40 boundary = 592 x 1104, 3000 steps, rate 4 ms):
  the objective + gradient and the Gauss-Newton product of one shot, bit-exact
  against the oracle's FP32 mode, each stated against FP64; two FP64 L-BFGS
  iterations driven by the GPU gradient equal the same iterations driven by
  the oracle's gradient (on a 600-step record, to bound the oracle's time).
- cfg3 (3D, 460 x 572 x 572): 12 absorbing forward levels (one and two levels
  per launch) and 8 fused backward levels of both modes (gradient, RTM) from
  a nonzero state, bit-exact against the oracle's reference arithmetic.
- cfg4 (256 shots, nz 512, apertures of 300-900 columns of a 4096-wide
  model, 2500 steps): the batched gradient equals the one-shot driver on each
  cropped model bit for bit (FP32 snapshots); FP16/BF16 snapshots stated
  against the FP64 oracle on two of the shots.
- cfg5 (served stream): 2D gradient requests of nz 256-512 / nx 256-1024 and
  3D modeling requests of 128-400 per dimension, 1000-3000 steps, submitted
  together to a 4-slot server: each result equals the same call alone.
"""
import numpy as np
import pytest

from common3d import rel
from vtiamd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------------------
# cfg2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg2(vti, ctx):
    P = W.cfg2()
    P["ctrue"] = W.cij(*P["true"])
    P["cinit"] = W.cij(*P["init"])
    P["taper"] = vti.init_taper2d(P["nx"], P["nz"], P["npad"])
    P["sl"] = P["sloc"][16:17].copy()
    _say("cfg2: observed data of the true model, shot 16")
    P["obs"] = vti.synthetic2d(ctx, *P["ctrue"], P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))
    return P


def _a2(P, nt=None):
    return (P["nx"], P["nz"], nt or P["nt"], P["npad"], P["dx"], P["dz"], P["dt"], P["rate"])


def test_cfg2_gradient_bitexact(vti, orc, ctx, cfg2):
    P = cfg2
    _say("cfg2: gradient, one shot, GPU")
    f, g = vti.objfunc_gradient2d(ctx, P["obs"], *P["cinit"], P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))
    _say("cfg2: gradient, oracle FP32")
    fo, go = orc.fg2d(orc.F32, False, P["obs"], *P["cinit"], None, P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))
    _say(f"cfg2 gradient: f_gpu={f:.9e} f_oracle={fo:.9e}")
    assert np.abs(go).max() > 0
    assert np.float32(f) == np.float32(fo)
    assert np.array_equal(g, go)


def test_cfg2_gradient_vs_fp64(vti, orc, ctx, cfg2):
    P = cfg2
    _say("cfg2: gradient, GPU and oracle FP64")
    f, g = vti.objfunc_gradient2d(ctx, P["obs"], *P["cinit"], P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))
    f64, g64 = orc.fg2d(orc.F64, False, P["obs"], *P["cinit"], None, P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))
    eg, ef = rel(g, g64), abs(f - f64) / abs(f64)
    _say(f"cfg2 gradient vs FP64: {eg:.2e}, objective {ef:.2e}")
    # the FP32 arithmetic of the reference (which the GPU reproduces bit for
    # bit, test_cfg2_gradient_bitexact) is 1.2e-4 / 2.2e-5 from FP64 here
    assert eg < 3e-4 and ef < 5e-5


def test_cfg2_gn_hessian_bitexact(vti, orc, ctx, cfg2):
    """Hv for a seeded random, smooth v-perturbation (dVEpsDel2Cij to dc)."""
    P = cfg2
    rng = np.random.default_rng(12)
    nnz, nnx = P["cinit"][0].shape
    dv = np.zeros((nnz, nnx), np.float32)
    dv[P["npad"]:-P["npad"], P["npad"]:-P["npad"]] = rng.standard_normal((P["nz"], P["nx"]))
    dv = (W._gauss_smooth(dv, 6.0) * 200.0).astype(np.float32)
    v, e, d = P["init"]
    dc = vti.dvepsdel_to_dcij(v, e, d, dv, np.zeros_like(dv), np.zeros_like(dv))[3:]
    _say("cfg2: GN Hv, GPU")
    h = vti.gn_hessian2d(ctx, *P["cinit"], *dc, P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))
    _say("cfg2: GN Hv, oracle FP32")
    ho = orc.fg2d(orc.F32, True, None, *P["cinit"], dc, P["w"], P["sl"], P["rloc"], P["taper"], *_a2(P))[1]
    assert np.abs(ho).max() > 0
    assert np.array_equal(h, ho)


def test_cfg2_lbfgs_two_iterations(vti, orc, ctx, cfg2):
    """Two iterations of the FP64 L-BFGS (m = 5) as fwi-lbfgs-c-param.cu runs
    them (parameter vepsdel: v/v0, eps/eps0, del/del0; checkEpsDel projects
    eps and del in place each evaluation), one shot, 600-step record, both
    with the host L-BFGS and with its device-resident form.  The objective
    is scaled to 1e6 at the initial model (the reference's stopping test
    |g|/max(1,|x|) <= 1e-5 is absolute, and the raw misfit of this record,
    about 1e-9, would end the run after its first step).  The
    GPU-driven run (objFuncGradientVEpsDel on the GPU) and the oracle-driven
    run (checkEpsDel, VEpsDel2Cij, the oracle's FP32 gradient,
    GradCij2GradVEpsDel, each pinned on its own) visit the same models bit
    for bit and decrease the objective."""
    P = cfg2
    nt = 601
    w = P["w"][:nt]
    _say("cfg2 L-BFGS: observed data, 600 steps")
    obs = vti.synthetic2d(ctx, *P["ctrue"], w, P["sl"], P["rloc"], P["taper"], *_a2(P, nt))
    v0, eps0, del0 = 100.0, 0.01, 0.01
    v, e, d = P["init"]
    nn = v.size
    shape = v.shape

    def run(grad, device):
        x = np.concatenate([(v / np.float32(v0)).ravel(), (e / np.float32(eps0)).ravel(),
                            (d / np.float32(del0)).ravel()]).astype(np.float32)
        opt = vti.LBFGS(x.size, m=5, ctx=ctx if device else None)
        fs, xs, iflag, f0 = [], [], 0, None
        while True:
            f, g, e2, d2 = grad(x[:nn], x[nn:2 * nn], x[2 * nn:])
            x[nn:2 * nn], x[2 * nn:] = e2.ravel(), d2.ravel()
            f0 = f if f0 is None else f0
            fs.append(f)
            iflag = opt.step(x, f / f0 * 1e6, (np.asarray(g, np.float64) / f0 * 1e6).astype(np.float32).ravel(),
                             iflag)
            xs.append(x.copy())
            if iflag != 1 or opt.iter >= 3 or len(fs) >= 24:
                break
        it = opt.iter
        opt.close()
        return fs, xs, it

    def gpu(xv, xe, xd):
        f, g, e2, d2, _ = vti.objfunc_gradient2d_vepsdel(ctx, obs, xv, xe, xd, v0, eps0, del0, w, P["sl"],
                                                         P["rloc"], P["taper"], *_a2(P, nt))
        return f, g, e2, d2

    def oracle(xv, xe, xd):
        ce, cd, _, _ = vti.check_eps_del(xe, xd, eps0, del0)
        c = [a.reshape(shape) for a in vti.vepsdel_to_cij(xv, ce, cd, v0, eps0, del0)]
        f, gc = orc.fg2d(orc.F32, False, obs, *c, None, w, P["sl"], P["rloc"], P["taper"], *_a2(P, nt))
        gg = vti.gradcij_to_gradvepsdel(gc[0], gc[1], gc[2], xv, ce, cd, v0, eps0, del0)
        rows = P["npad"] + 1  # wbottom 0: the reference zeroes npad + 1 rows
        g = np.stack([a.reshape(shape) for a in gg])
        g[:, :rows] = 0
        return float(np.float32(f)), g, ce, cd

    for device in (False, True):
        _say(f"cfg2 L-BFGS ({'device' if device else 'host'} vectors): GPU-driven")
        fg, xg, itg = run(gpu, device)
        _say(f"  objectives {fg}, iterations {itg}")
        _say("cfg2 L-BFGS: oracle-driven")
        fo, xo, ito = run(oracle, device)
        assert itg >= 2 and itg == ito
        assert fg == fo
        for a, b in zip(xg, xo):
            assert np.array_equal(a, b)
        assert fg[-1] < fg[0]


# ---------------------------------------------------------------------------
# cfg3
# ---------------------------------------------------------------------------
NZ3, NY3, NX3, NPAD3, D3, DT3 = 460, 572, 572, 30, 12.5, 1e-3


@pytest.fixture(scope="module")
def cfg3():
    c11, c13, c33, _ = W.layered3d(NZ3, NY3, NX3)
    return c11, c13, c33


def _state(seed, shape, scale=1e-3):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal(shape, dtype=np.float32) * np.float32(scale)) for _ in range(4)]


def _damp(vti, shape, npad):
    """The absorbing factor d = dxy * dz (float) of every point the steps
    update, 1 elsewhere.  The reference damps the current level in place after
    each block's step (abc3d.cu:82-125), so the level it later reads as prev
    carries a second factor; the propagator stores that level damped once and
    applies the second factor when it reads it (the state it holds is prev /
    d there).  So the oracle starts from prev * d and its final prev equals
    the propagator's times d."""
    nz, ny, nx = shape
    d = vti.init_damping3d(nx, ny, nz, npad)
    dd = (d[: nx * ny].reshape(ny, nx)[None] * d[nx * ny:][:, None, None]).astype(np.float32)
    inner = np.zeros(shape, bool)
    inner[: 4 * (nz // 4), 4:ny - 4, 4:nx - 4] = True
    return np.where(inner, dd, np.float32(1)).astype(np.float32)


@pytest.mark.parametrize("levels", [1, 2])
def test_cfg3_forward_window_bitexact(vti, orc, ctx, cfg3, levels):
    c = cfg3
    shape = c[0].shape
    cx, px, cz, pz = _state(31, shape)
    _say(f"cfg3: 12 absorbing forward levels, {levels} per launch, 460x572x572")
    with vti.Prop3D(ctx, NX3, NY3, NZ3, NPAD3, D3, D3, D3, DT3) as p:
        p.set_model(*c)
        p.set_levels(levels)
        p.set_state(cur_x=cx, cur_z=cz, prev_x=px, prev_z=pz)
        p.forward_steps(12, absorbing=True)
        got = p.get_state()
    _say("cfg3: oracle, reference arithmetic")
    dd = _damp(vti, shape, NPAD3)
    st = [cx, px * dd, cz, pz * dd]
    orc.steps3d(orc.REF, st, *c, NPAD3, D3, D3, D3, DT3, 12, absorbing=True)
    # 12 levels (even): the oracle's arrays hold (cur, prev) in their original roles
    assert np.abs(st[0]).max() > 0
    gx, gz, gpx, gpz = got
    for name, a, b in (("cur_x", gx, st[0]), ("cur_z", gz, st[2]), ("prev_x", gpx * dd, st[1]),
                       ("prev_z", gpz * dd, st[3])):
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{name}: {len(bad)} points differ, first {bad[:3].tolist()}"


@pytest.mark.parametrize("mode", [0, 1])
def test_cfg3_backward_window_bitexact(vti, orc, ctx, cfg3, mode):
    c = cfg3
    shape = c[0].shape
    sig = _state(41, shape)
    lam = _state(42, shape, 1e-6)
    _say(f"cfg3: 8 fused backward levels, mode {mode}, 460x572x572")
    with vti.Prop3D(ctx, NX3, NY3, NZ3, NPAD3, D3, D3, D3, DT3) as p:
        p.set_model(*c)
        p.set_state(cur_x=sig[0], cur_z=sig[2], prev_x=sig[1], prev_z=sig[3])
        p.set_adjoint(mode, cur_x=lam[0], cur_z=lam[2], prev_x=lam[1], prev_z=lam[3])
        p.backward_steps(8, mode=mode)
        s_gpu = p.get_state()
        l_gpu = p.get_adjoint(mode)
    _say("cfg3: oracle, reference arithmetic")
    # the adjoint side is damped (prev stored damped once less, see _damp);
    # the source-side reconstruction is not
    dd = _damp(vti, shape, NPAD3)
    lam = [lam[0], lam[1] * dd, lam[2], lam[3] * dd]
    acc = [np.zeros(shape, np.float32) for _ in range(3 if mode == 0 else 1)]
    sig_o, lam_o = orc.bwd_steps3d(mode, 1, sig, lam, acc, *c, NPAD3, D3, D3, D3, DT3, 8)
    assert np.abs(acc[0]).max() > 0

    def same(name, a, b):
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{name}: {len(bad)} points differ, first {bad[:3].tolist()}"

    for k, (a, b) in enumerate(zip(s_gpu, (sig_o[0], sig_o[2], sig_o[1], sig_o[3]))):
        same(f"sigma[{k}]", a, b)
    lg = list(l_gpu[:4])
    lg[2], lg[3] = lg[2] * dd, lg[3] * dd
    for k, (a, b) in enumerate(zip(lg, (lam_o[0], lam_o[2], lam_o[1], lam_o[3]))):
        same(f"lambda[{k}]", a, b)
    for k, (a, b) in enumerate(zip(l_gpu[4], acc)):
        same(f"accumulator[{k}]", a, b)


# ---------------------------------------------------------------------------
# cfg4
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg4(vti, ctx):
    B = W.cfg4()
    B["ct"] = W.cij(*B["true"])
    B["c"] = W.cij(*B["init"])
    nnt = (B["nt"] - 1) // 4 + 1
    _say("cfg4: observed data of 256 shots (one-shot modeling calls)")
    out = []
    for c0, nx, ix, iz, nr, r0 in B["shots"]:
        wd = nx + 2 * B["npad"]
        taper = vti.init_taper2d(nx, B["nz"], B["npad"])
        d = vti.synthetic2d(ctx, *[np.ascontiguousarray(a[:, c0:c0 + wd]) for a in B["ct"]], B["w"],
                            np.array([[ix, iz, nr, 0]], np.int32), B["rloc"][r0:r0 + nr], taper, nx, B["nz"], B["nt"],
                            B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
        assert d.shape == (nnt, nr)
        out.append(d.ravel())
    B["data"] = np.concatenate(out)
    return B


def _plain4(vti, ctx, B, idx):
    g = np.zeros((3,) + B["c"][0].shape, np.float64)
    f = 0.0
    nnt = (B["nt"] - 1) // 4 + 1
    offs = np.concatenate([[0], np.cumsum(B["shots"][:, 4] * nnt)])
    for k in idx:
        c0, nx, ix, iz, nr, r0 = B["shots"][k]
        wd = nx + 2 * B["npad"]
        taper = vti.init_taper2d(nx, B["nz"], B["npad"])
        d = B["data"][offs[k]:offs[k + 1]].reshape(nnt, nr)
        fs, gs = vti.objfunc_gradient2d(ctx, d, *[np.ascontiguousarray(a[:, c0:c0 + wd]) for a in B["c"]], B["w"],
                                        np.array([[ix, iz, nr, 0]], np.int32), B["rloc"][r0:r0 + nr], taper, nx,
                                        B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
        f += fs
        g[:, :, c0:c0 + wd] += gs.astype(np.float64)
    return f, g


def _batch_args(B):
    return (B["NX"], B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])


def test_cfg4_batch_equals_one_shot_driver(vti, ctx, cfg4):
    B = cfg4
    _say("cfg4: batched gradient, 256 shots, tuned launch shape")
    shape = vti.tune_batch2d(ctx, B["shots"], B["rloc"], *_batch_args(B))
    _say(f"cfg4: launch shape {shape}")
    fb, gb = vti.objfunc_gradient2d_batch(ctx, B["data"], *B["c"], B["w"], B["shots"], B["rloc"], *_batch_args(B),
                                          max_wave=shape["max_wave"], levels=shape["levels"],
                                          tile_rows=shape["tile_rows"])
    _say("cfg4: one-shot driver on each cropped model")
    fp, gp = _plain4(vti, ctx, B, range(len(B["shots"])))
    _say(f"cfg4: f batch {fb:.9e} one-shot {fp:.9e}")
    assert np.abs(gp).max() > 0
    assert np.array_equal(gb, gp)


@pytest.fixture(scope="module")
def cfg4_fp64(orc, vti, cfg4):
    """The FP64 oracle's gradient of the two narrowest of the 256 shots, each
    on its cropped model, summed in FP64 on the global grid."""
    B = cfg4
    sel = [int(k) for k in np.argsort(B["shots"][:, 1], kind="stable")[:2]]
    nnt = (B["nt"] - 1) // 4 + 1
    offs = np.concatenate([[0], np.cumsum(B["shots"][:, 4] * nnt)])
    _say(f"cfg4: oracle FP64 on shots {sel} (cropped models)")
    g64 = np.zeros((3,) + B["c"][0].shape, np.float64)
    for k in sel:
        c0, nx, ix, iz, nr, r0 = B["shots"][k]
        wd = nx + 2 * B["npad"]
        taper = vti.init_taper2d(nx, B["nz"], B["npad"])
        d = B["data"][offs[k]:offs[k + 1]].reshape(nnt, nr)
        _, go = orc.fg2d(orc.F64, False, d, *[np.ascontiguousarray(a[:, c0:c0 + wd]) for a in B["c"]], None, B["w"],
                         np.array([[ix, iz, nr, 0]], np.int32), B["rloc"][r0:r0 + nr], taper, nx, B["nz"], B["nt"],
                         B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
        g64[:, :, c0:c0 + wd] += go
    return sel, g64


@pytest.mark.parametrize("snap", ["fp16", "bf16", "s16"])
def test_cfg4_16bit_snapshots_vs_fp64(vti, ctx, cfg4, cfg4_fp64, snap):
    """Two of the 256 shots with 16-bit snapshots against the FP64 oracle."""
    B = cfg4
    sel, g64 = cfg4_fp64
    nnt = (B["nt"] - 1) // 4 + 1
    offs = np.concatenate([[0], np.cumsum(B["shots"][:, 4] * nnt)])
    shots = B["shots"][sel].copy()
    rl, data = [], []
    for j, k in enumerate(sel):
        r0, nr = B["shots"][k][5], B["shots"][k][4]
        shots[j][5] = len(rl)
        rl += list(map(tuple, B["rloc"][r0:r0 + nr]))
        data.append(B["data"][offs[k]:offs[k + 1]])
    rl, data = np.array(rl, np.int32), np.concatenate(data)
    _say(f"cfg4: shots {sel} with {snap} snapshots")
    _, g16 = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], shots, rl, *_batch_args(B), snap=snap)
    _, g32 = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], shots, rl, *_batch_args(B))
    e16, e32 = rel(g16, g64), rel(g32, g64)
    _say(f"cfg4 {snap}: gradient vs FP64 {e16:.2e} (FP32 snapshots {e32:.2e})")
    assert e32 < 1e-4
    assert e16 < {"fp16": 3e-4, "bf16": 3e-3, "s16": 1e-4}[snap]


# ---------------------------------------------------------------------------
# cfg5
# ---------------------------------------------------------------------------
def test_cfg5_served_requests_equal_alone(vti, ctx):
    """2D gradient requests (nz x nx = 384 x 800, 256 x 1024, 512 x 300 +
    40 boundary; 1500, 1000, 2000 steps; two of them twice, so that merging
    applies) and 3D modeling requests (200 x 180 x 160 at 1000 steps, 128 x
    400 x 256 at 1200) submitted together to a 4-slot server."""
    npad, dx, dt, rate = 40, 10.0, 1e-3, 4e-3
    jobs = []
    with vti.Server(0, slots=4) as srv:
        for (nz, nx, nt, seed) in ((384, 800, 1501, 1), (256, 1024, 1001, 2), (512, 300, 2001, 3),
                                   (384, 800, 1501, 4), (256, 1024, 1001, 5)):
            true, init = W.fwi2d_model(nx, nz, npad, seed % 3, 8, smooth=8.0)
            c, ct = W.cij(*init), W.cij(*true)
            taper = vti.init_taper2d(nx, nz, npad)
            w = W.ricker(nt, dt, 12.0, 0.1)
            sl = np.array([[npad + nx // 3 + 17 * seed, npad + 2, nx, 0]], np.int32)
            rl = np.array([[npad + i, npad + 1] for i in range(nx)], np.int32)
            a = (nx, nz, nt, npad, dx, dx, dt, rate)
            obs = vti.synthetic2d(ctx, *ct, w, sl, rl, taper, *a)
            _say(f"cfg5: 2D request {nz}x{nx}, {nt - 1} steps")
            h = srv.put_model(*c)
            t = srv.grad2d(h, obs, w, sl, rl, taper, *a)
            jobs.append(("2d", t, (obs, c, w, sl, rl, taper, a)))
        for (g, nt) in (((200, 180, 160), 1000), ((128, 400, 256), 1200)):
            c11, c13, c33, _ = W.layered3d(*g, seed=sum(g))
            nz, ny, nx = g
            d = 12.5
            rec = np.array([(ix * d + 0.1, iy * d + 0.1, 32 * d + 0.1, 1.0) for iy in (ny // 3, 2 * ny // 3)
                            for ix in range(30, nx - 30, 4)], np.float32)
            sou = np.array([[nx // 2 * d + 0.1, ny // 2 * d + 0.1, 34 * d + 0.1, len(rec), 0]], np.float32)
            w = W.ricker(nt, dt, 12.0, 0.1)
            _say(f"cfg5: 3D request {g}, {nt} steps")
            h = srv.put_model(c11, c13, c33)
            t = srv.model3d(h, sou, np.arange(1), rec, w, g, nt, 30, (0.0, 0.0, 0.0), (d, d, d), dt, rate)
            jobs.append(("3d", t, (sou, rec, w, (c11, c13, c33), nt)))
        for kind, t, args in jobs:
            res, _, _ = srv.wait(t)
            _say(f"cfg5: {kind} request done, comparing with the call alone")
            if kind == "2d":
                obs, c, w, sl, rl, taper, a = args
                f, g = vti.objfunc_gradient2d(ctx, obs, *c, w, sl, rl, taper, *a)
                assert res[0] == np.float32(f)
                assert np.array_equal(res[1], g)
            else:
                sou, rec, w, c, nt = args
                dd = vti.model_data3d(ctx, sou, np.arange(1), rec, w, *c, nt, 30, (0.0, 0.0, 0.0),
                                      (12.5, 12.5, 12.5), dt, rate)
                assert np.array_equal(res, dd)
        st = srv.stats()
    _say(f"cfg5: merged runs {st['merged_runs']}, requests in them {st['merged_requests']}")
