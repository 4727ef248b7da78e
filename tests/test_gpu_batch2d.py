"""Batched 2D gradient (many shots on aperture sub-grids of one model, one
or two time levels per launch) against the single-shot driver on the cropped
models: per shot the same FP32 gradient, summed in FP64 in shot order -- so
the global gradients are equal bit for bit for every launch shape.  With
16-bit snapshots the gradient is stated against the FP64 oracle."""
import numpy as np
import pytest

from common2d import layered2d
from common3d import cij_from, rel, ricker

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def batch_problem(vti, NX=160, nz=40, npad=10, nt=161, ns=5, seed=8, nx_range=(40, 90), sz=2, rz=1):
    """ns shots on random apertures nx_range wide of an NX-wide model; source
    sz and receivers rz rows below the top padding."""
    rng = np.random.default_rng(seed)
    NNX, nnz = NX + 2 * npad, nz + 2 * npad
    v, e, d = layered2d(nnz, NNX, seed=seed)
    c = cij_from(v, e, d)
    vt = v.copy()
    vt[nnz // 2 - 4: nnz // 2 + 4, :] *= 1.04
    ct = cij_from(vt, e, d)
    shots, rloc = [], []
    for s in range(ns):
        nx = int(rng.integers(*nx_range))
        c0 = int(rng.integers(0, NX - nx + 1))
        w = nx + 2 * npad
        step = int(rng.integers(1, 4))
        rx = list(range(npad, npad + nx, step))
        shots.append((c0, nx, npad + int(rng.integers(4, nx - 4)), npad + sz, len(rx), len(rloc)))
        rloc += [(x, npad + rz) for x in rx]
    return dict(NX=NX, nz=nz, npad=npad, NNX=NNX, nnz=nnz, nt=nt, dx=10.0, dz=10.0, dt=1e-3, rate=4e-3, c=c, ct=ct,
                w=ricker(nt, 1e-3, 20.0, 0.05), shots=np.array(shots, np.int32), rloc=np.array(rloc, np.int32))


def crop(a, c0, w):
    return np.ascontiguousarray(a[:, c0:c0 + w])


def plain(vti, ctx, B, data):
    """The single-shot driver on each shot's cropped model, summed in FP64 in shot order."""
    g = np.zeros((3, B["nnz"], B["NNX"]), np.float64)
    f = 0.0
    off = 0
    nnt = (B["nt"] - 1) // 4 + 1
    for c0, nx, ix, iz, nr, r0 in B["shots"]:
        w = nx + 2 * B["npad"]
        taper = vti.init_taper2d(nx, B["nz"], B["npad"])
        sl = np.array([[ix, iz, nr, 0]], np.int32)
        rl = B["rloc"][r0:r0 + nr]
        d = data[off:off + nnt * nr].reshape(nnt, nr)
        off += nnt * nr
        fs, gs = vti.objfunc_gradient2d(ctx, d, *[crop(c, c0, w) for c in B["c"]], B["w"], sl, rl, taper, nx, B["nz"],
                                        B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
        f += fs
        g[:, :, c0:c0 + w] += gs.astype(np.float64)
    return f, g


def observed(vti, ctx, B):
    out = []
    for c0, nx, ix, iz, nr, r0 in B["shots"]:
        w = nx + 2 * B["npad"]
        taper = vti.init_taper2d(nx, B["nz"], B["npad"])
        d = vti.synthetic2d(ctx, *[crop(c, c0, w) for c in B["ct"]], B["w"], np.array([[ix, iz, nr, 0]], np.int32),
                            B["rloc"][r0:r0 + nr], taper, nx, B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"],
                            B["rate"])
        out.append(d.ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("max_wave", [0, 2])
def test_batch_equals_single_shot_driver(vti, ctx, max_wave):
    B = batch_problem(vti)
    _say("batch2d", len(B["shots"]), "shots, max_wave", max_wave)
    data = observed(vti, ctx, B)
    fb, gb = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"], B["nz"],
                                          B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"], max_wave=max_wave)
    fp, gp = plain(vti, ctx, B, data)
    assert np.abs(gp).max() > 0
    assert np.array_equal(gb, gp)
    assert abs(fb - fp) <= 1e-6 * fp


def test_batch_rejects_subgrid_outside_model(vti, ctx):
    B = batch_problem(vti, ns=1)
    sh = B["shots"].copy()
    sh[0, 0] = B["NX"]  # starts past the last column that leaves room for it
    with pytest.raises(vti.VtiError, match="outside the model"):
        vti.objfunc_gradient2d_batch(ctx, np.zeros(10**5, np.float32), *B["c"], B["w"], sh, B["rloc"], B["NX"],
                                     B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])


@pytest.mark.parametrize("levels,rows", [(1, 8), (2, 4), (2, 8), (2, 16)])
def test_launch_shapes_bit_exact(vti, ctx, levels, rows):
    """Two levels per launch (temporal blocking) and every tile height give
    the one-shot driver's bits; nt = 4k+1 leaves an odd level count, so each
    direction also runs one single-level launch."""
    B = batch_problem(vti, seed=11, ns=4)
    _say("batch2d levels", levels, "rows", rows)
    data = observed(vti, ctx, B)
    fb, gb = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"], B["nz"],
                                          B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"], levels=levels,
                                          tile_rows=rows)
    fp, gp = plain(vti, ctx, B, data)
    assert np.abs(gp).max() > 0
    assert np.array_equal(gb, gp)
    assert abs(fb - fp) <= 1e-6 * fp


@pytest.mark.parametrize("rows,snap", [(4, "fp32"), (8, "fp32"), (16, "fp32"), (4, "s16"), (8, "s16"), (16, "s16"),
                                       (16, "bf16")])
def test_inner_tiles_bit_exact(vti, ctx, rows, snap):
    """Apertures wide and deep enough that most two-level tiles lie where the
    taper is 1 (the kernels' inner-tile form: no clamps, bounds tests or
    taper factors), with the source and the receiver row inside such tiles:
    the one-shot driver's bits with FP32 snapshots, and with 16-bit ones the
    bits of single-level launches, which have no inner-tile form."""
    B = batch_problem(vti, NX=600, nz=120, npad=20, nt=121, ns=4, seed=15, nx_range=(150, 300), sz=40, rz=30)
    _say("batch2d inner tiles rows", rows, snap)
    # half the observed data: a residual that is not zero within the short run
    data = observed(vti, ctx, B) * np.float32(0.5)
    run = lambda **kw: vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"],
                                                    B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"],
                                                    snap=snap, **kw)
    fb, gb = run(levels=2, tile_rows=rows, max_wave=3)
    assert np.abs(gb).max() > 0
    if snap == "fp32":
        fp, gp = plain(vti, ctx, B, data)
        assert np.array_equal(gb, gp)
        assert abs(fb - fp) <= 1e-6 * fp
    else:
        f1, g1 = run(levels=1, tile_rows=8)
        assert np.array_equal(gb, g1) and fb == f1


def oracle_f64(orc, vti, B, data):
    """The FP64 oracle of objFuncGradient_f on each shot's cropped model, summed."""
    g = np.zeros((3, B["nnz"], B["NNX"]), np.float64)
    off = 0
    nnt = (B["nt"] - 1) // 4 + 1
    for c0, nx, ix, iz, nr, r0 in B["shots"]:
        w = nx + 2 * B["npad"]
        taper = vti.init_taper2d(nx, B["nz"], B["npad"])
        d = data[off:off + nnt * nr]
        off += nnt * nr
        _, gs = orc.fg2d(orc.F64, False, d, *[crop(c, c0, w) for c in B["c"]], None, B["w"],
                         np.array([[ix, iz, nr, 0]], np.int32), B["rloc"][r0:r0 + nr], taper, nx, B["nz"], B["nt"],
                         B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
        g[:, :, c0:c0 + w] += gs
    return g


# Measured on MI355X (zero data, 5 shots, relative L2 of the summed gradient
# against the FP64 oracle): FP32 snapshots 1.4e-5 (the FP32 stencil itself),
# FP16 1.1e-4, BF16 9.8e-4 -- the formats' unit roundoffs 2^-11 and 2^-8
# scaled down by averaging over the imaging sum.  S16 (16-bit integers with a
# power-of-two scale per 64-point segment) keeps 15 bits relative to each
# segment's largest value and has to stay within the 1e-4 bar.
@pytest.mark.parametrize("snap,tol", [("fp16", 3e-4), ("bf16", 3e-3), ("s16", 1e-4)])
def test_16bit_snapshots_vs_fp64_oracle(vti, orc, ctx, snap, tol):
    B = batch_problem(vti, seed=12)
    nnt = (B["nt"] - 1) // 4 + 1
    data = np.zeros(int(sum(nnt * s[4] for s in B["shots"])), np.float32)
    g64 = oracle_f64(orc, vti, B, data)
    _, g32 = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"], B["nz"],
                                          B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
    _, g16 = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"], B["nz"],
                                          B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"], snap=snap)
    e32, e16 = rel(g32, g64), rel(g16, g64)
    _say(f"batch2d snapshots fp32 rel vs FP64 {e32:.3e}, {snap} {e16:.3e}")
    assert e32 < 1e-4
    assert e16 < tol
    # a 16-bit run with two-level launches equals one with single-level launches
    _, g16b = vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"], B["nz"],
                                           B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"], snap=snap,
                                           levels=1)
    assert np.array_equal(g16, g16b)


def test_autotuned_launch_shape(vti, ctx):
    """The tuner returns one of its candidates, caches it, and the batch run
    with it gives the same bits as the default launch shape."""
    B = batch_problem(vti, seed=13, ns=5)
    a = (B["shots"], B["rloc"], B["NX"], B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
    t = vti.tune_batch2d(ctx, *a)
    _say("batch2d tuned", t)
    # the wave size is modelled from the Infinity Cache: these small shots all fit
    assert t["levels"] in (1, 2) and t["tile_rows"] in (8, 16) and t["max_wave"] == 0
    assert t["trial_ms"] > 0
    assert vti.tune_batch2d(ctx, *a) == t  # cached
    data = observed(vti, ctx, B)
    run = lambda **kw: vti.objfunc_gradient2d_batch(ctx, data, *B["c"], B["w"], B["shots"], B["rloc"], B["NX"],
                                                    B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"],
                                                    **kw)
    f0, g0 = run()
    f1, g1 = run(levels=t["levels"], tile_rows=t["tile_rows"], max_wave=t["max_wave"])
    assert np.array_equal(g0, g1) and f0 == f1


@pytest.mark.parametrize("snap,levels,max_wave",
                         [("fp32", 2, 0), ("fp32", 1, 3), ("bf16", 2, 2), ("s16", 2, 1), ("fp16", 1, 3)])
def test_graph_plan_replays_equal_direct_calls(vti, ctx, snap, levels, max_wave):
    """A captured plan (hipGraph of the whole launch sequence) replayed on
    three models, with other drivers' work and synchronisations between the
    replays, gives each time the direct batched call's bits on that model --
    also the third replay, back on the first model.  With max_wave > 0 the
    plan holds several waves, two graphs each, replayed two at a time on two
    streams through two alternating sets of wavefield, snapshot and trace
    buffers."""
    B = batch_problem(vti, seed=14, ns=4)
    _say("batch2d plan", snap, "levels", levels, "max_wave", max_wave)
    data = observed(vti, ctx, B)
    a = (B["NX"], B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
    models = [B["c"], tuple((c * np.float32(1.01)).astype(np.float32) for c in B["c"]), B["c"]]
    with vti.BatchPlan(ctx, data, B["w"], B["shots"], B["rloc"], *a, max_wave=max_wave, levels=levels,
                       snap=snap) as plan:
        for k, m in enumerate(models):
            fg, gg = plan.run(*m)
            fd, gd = vti.objfunc_gradient2d_batch(ctx, data, *m, B["w"], B["shots"], B["rloc"], *a,
                                                  max_wave=max_wave, levels=levels, snap=snap)
            _say("  replay", k, fg, fd)
            assert np.abs(gd).max() > 0
            assert np.array_equal(gg, gd) and fg == fd
            # eager work of another driver between replays
            vti.objfunc_gradient2d_batch(ctx, data[::-1].copy(), *B["ct"], B["w"], B["shots"], B["rloc"], *a,
                                         max_wave=2)


@pytest.mark.parametrize("snap,tol", [("s16", 1e-4), ("bf16", 3e-3)])
def test_context_snapshot_format_one_shot_and_served(vti, orc, ctx, snap, tol):
    """A context set to 16-bit snapshots runs a one-shot objFuncGradient_f
    through the batched driver with that storage: its gradient is the batched
    call's on the same shot bit for bit, within the format's bar of the FP64
    oracle, and a server set to the same format returns the same bits; the
    default context keeps FP32 (the one-shot loop's bits)."""
    B = batch_problem(vti, seed=16, ns=1)
    data = observed(vti, ctx, B) * np.float32(0.5)
    nnt = (B["nt"] - 1) // 4 + 1
    c0, nx, ix, iz, nr, r0 = (int(q) for q in B["shots"][0])
    w = nx + 2 * B["npad"]
    cc = [crop(c, c0, w) for c in B["c"]]
    taper = vti.init_taper2d(nx, B["nz"], B["npad"])
    sl = np.array([[ix, iz, nr, 0]], np.int32)
    rl = B["rloc"][r0:r0 + nr]
    d = data[:nnt * nr].reshape(nnt, nr)
    geo = (nx, B["nz"], B["nt"], B["npad"], B["dx"], B["dz"], B["dt"], B["rate"])
    _say("context snapshots", snap, "shot width", w)
    f32, g32 = vti.objfunc_gradient2d(ctx, d, *cc, B["w"], sl, rl, taper, *geo)
    with vti.Context(0) as c16:
        c16.set_snapshots(snap)
        f16, g16 = vti.objfunc_gradient2d(c16, d, *cc, B["w"], sl, rl, taper, *geo)
        c16.set_snapshots("fp32")
        f32b, g32b = vti.objfunc_gradient2d(c16, d, *cc, B["w"], sl, rl, taper, *geo)
    assert np.array_equal(g32, g32b) and f32 == f32b
    assert not np.array_equal(g16, g32)
    shot1 = np.array([[0, nx, ix, iz, nr, 0]], np.int32)
    fb, gb = vti.objfunc_gradient2d_batch(ctx, d.ravel(), *cc, B["w"], shot1, rl, *geo, snap=snap)
    assert np.array_equal(g16, gb.astype(np.float32))
    assert abs(float(f16) - fb) <= 1e-6 * fb
    _, g64 = orc.fg2d(orc.F64, False, d.ravel(), *cc, None, B["w"], sl, rl, taper, *geo)
    e16, e32 = rel(g16, g64), rel(g32, g64)
    _say(f"  rel vs FP64: {snap} {e16:.3e}, fp32 {e32:.3e}")
    assert e16 < tol
    with vti.Server(0, slots=2, snap=snap) as srv:
        h = srv.put_model(*cc)
        t = srv.grad2d(h, d, B["w"], sl, rl, taper, *geo)
        (fs, gs), _, _ = srv.wait(t)
    assert np.array_equal(gs, g16) and np.float32(fs) == np.float32(f16)
