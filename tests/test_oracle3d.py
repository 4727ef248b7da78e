"""Pins of the 3D oracle (CPU): what the mathematics fixes, independently of
the oracle's own code.

- one step on separable sinusoids against the closed-form eigenvalue of the
  8th-order stencil, with spatially random c11/c13/c33 (a transposed
  coefficient, a wrong scaling or a swapped field fails);
- the reference's block-sequential damping order restated independently in
  numpy, against the oracle's block-by-block emulation of it;
- time reversibility of the undamped step (the random-boundary
  reconstruction relies on it) and of the gradient kernel's form;
- agreement of the FP32 / reference-arithmetic modes with FP64;
- linearity of the RTM image in the data and the step-count rule of the
  reference pipeline.
"""
import numpy as np
import pytest

from common3d import cij_from, layered_model, rel, ricker

C = np.array([-2.84722222222, 1.6, -0.2, 0.02539682539, -0.00178571428])


def lam(k):
    """Eigenvalue of the reference's 9-point second difference on exp(i k n)."""
    return C[0] + 2 * sum(C[m] * np.cos(m * k) for m in range(1, 5))


def _fields(nz, ny, nx, seed=0):
    rng = np.random.default_rng(seed)
    c11 = rng.uniform(4e6, 9e6, (nz, ny, nx))
    c33 = rng.uniform(2e6, 4e6, (nz, ny, nx))
    c13 = rng.uniform(1e6, 3e6, (nz, ny, nx))
    return c11, c13, c33


def test_lmax_rule(orc):
    # nround*ngpu*64 steps, nround = (nt-2)/(ngpu*64)  (objFuncGradient3d_f.cu:539)
    assert orc.lmax3d(3000, 1) == (2998 // 64) * 64 + 1
    assert orc.lmax3d(3010, 1) == 3009
    assert orc.lmax3d(3010, 2) == (3008 // 128) * 128 + 1
    assert orc.lmax3d(60, 1) == 1


@pytest.mark.parametrize("mode", ["f64", "ref"])
def test_single_step_closed_form(orc, mode):
    nz, ny, nx, npad = 24, 26, 28, 5
    dx, dy, dz, dt = 10.0, 12.0, 8.0, 1e-3
    kx, ky, kz = 0.7, 0.45, 1.1
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    sx = np.cos(kx * x) * np.cos(ky * y) * np.sin(kz * z + 0.3)
    sz = 0.6 * np.sin(kx * x + 0.2) * np.cos(ky * y + 0.1) * np.cos(kz * z)
    rng = np.random.default_rng(1)
    px, pz = rng.standard_normal((2, nz, ny, nx))
    c11, c13, c33 = _fields(nz, ny, nx)
    # closed form: the stencil eigenvalue on each separable factor, scaled by
    # the reference's float dt2/d2 factors (dt2dx2 etc. are float kernel arguments)
    f = np.float32
    dt2 = f(dt) * f(dt)
    sxs, sys_, szs = (float(dt2 / f(h * h)) for h in (dx, dy, dz))
    lxy_sx = (lam(kx) * sxs + lam(ky) * sys_) * sx
    lz_sz = lam(kz) * szs * sz
    want_x = 2 * sx - px + (c11 * lxy_sx + c13 * lz_sz)
    want_z = 2 * sz - pz + (c13 * lxy_sx + c33 * lz_sz)
    inner = (slice(4, nz - 4), slice(4, ny - 4), slice(4, nx - 4))
    if mode == "f64":
        st = [a.copy() for a in (sx, px, sz, pz)]
        orc.steps3d_f64(st, c11, c13, c33, npad, dx, dy, dz, dt, 1, absorbing=False)
        # after one level the new field sits in the array passed as prev
        got_x, got_z = st[1], st[3]
        tol = 1e-9
    else:
        st = [np.ascontiguousarray(a, np.float32) for a in (sx, px, sz, pz)]
        cs = [np.ascontiguousarray(a, np.float32) for a in (c11, c13, c33)]
        orc.steps3d(orc.REF, st, *cs, npad, dx, dy, dz, dt, 1, absorbing=False)
        got_x, got_z = st[1].astype(np.float64), st[3].astype(np.float64)
        tol = 3e-6
    # the scaled stencil uses float(dt2/dx2); compare to the exact-coefficient closed form
    assert rel(got_x[inner], want_x[inner]) < tol
    assert rel(got_z[inner], want_z[inner]) < tol
    # border columns are never updated (interior x,y in [4, n-4))
    assert np.array_equal(got_x[:, :, :4], np.asarray(st[1])[:, :, :4].astype(np.float64))
    p_in = px.astype(np.float32).astype(np.float64) if mode == "ref" else px
    assert np.array_equal(got_x[:, :, :4], p_in[:, :, :4])


def test_block_order_damping_rule(orc):
    """One absorbing step, FP64: the oracle's block-by-block run equals the rule
    'z-neighbours in the block above read the level damped once more'."""
    nz, ny, nx, npad = 20, 22, 24, 6
    dx, dt = 10.0, 1e-3
    rng = np.random.default_rng(4)
    sx, px, sz, pz = rng.standard_normal((4, nz, ny, nx))
    c11, c13, c33 = _fields(nz, ny, nx, seed=5)
    d = orc.init_damping3d(nx, ny, nz, npad)
    dxy = d[: nx * ny].reshape(ny, nx).astype(np.float64)
    dzv = d[nx * ny:].astype(np.float64)
    # the reference multiplies by the float product damping[x+y*nx]*damping[nxy+z]
    damp = (dxy[None].astype(np.float32) * dzv[:, None, None].astype(np.float32)).astype(np.float64)
    inner = np.zeros((nz, ny, nx), bool)
    inner[:, 4:ny - 4, 4:nx - 4] = True
    dt2 = float(np.float32(dt) * np.float32(dt))
    s2 = np.float32(dt2) / np.float32(dx * dx)
    s2 = float(s2)

    def d2(u, axis):
        out = C[0] * u
        for m in range(1, 5):
            out = out + C[m] * (np.roll(u, m, axis) + np.roll(u, -m, axis))
        return out

    want_x = np.array(px)
    want_z = np.array(pz)
    zmax = 4 * (nz // 4)
    for z in range(zmax):
        b = z // 4
        col = np.zeros((9, ny, nx))
        for k in range(-4, 5):
            zz = z + k
            if zz < 0 or zz >= zmax:
                continue
            v = sz[zz]
            if zz < 4 * b:  # block above: its current level was damped before this block ran
                v = v * np.where(inner[zz], damp[zz], 1.0)
            col[k + 4] = v
        tz = (C[0] * col[4] + sum(C[m] * (col[4 - m] + col[4 + m]) for m in range(1, 5))) * s2
        txy = (d2(sx[z], 1) + d2(sx[z], 0)) * s2
        dd = np.where(inner[z], damp[z], 1.0)
        # prev is stored already damped twice (as next, then as cur), so it is read as is
        nxv = c11[z] * txy + c13[z] * tz + 2 * sx[z] - px[z]
        nzv = c13[z] * txy + c33[z] * tz + 2 * sz[z] - pz[z]
        want_x[z] = np.where(inner[z], nxv * dd, px[z])
        want_z[z] = np.where(inner[z], nzv * dd, pz[z])
    st = [a.copy() for a in (sx, px, sz, pz)]
    orc.steps3d_f64(st, c11, c13, c33, npad, dx, dx, dx, dt, 1, absorbing=True)
    assert rel(st[1], want_x) < 1e-12
    assert rel(st[3], want_z) < 1e-12
    # the cur level left behind is damped once more in the interior
    assert rel(st[0], np.where(inner, sx * damp, sx)) < 1e-15


@pytest.mark.parametrize("gradient_form", [False, True])
def test_time_reversal_fp64(orc, gradient_form):
    nz, ny, nx, npad = 28, 24, 26, 6
    v, e, dl = layered_model(nz, ny, nx, seed=2)
    c11, c13, c33 = (a.astype(np.float64) for a in cij_from(v, e, dl))
    rng = np.random.default_rng(3)
    u0 = np.zeros((nz, ny, nx))
    u0[6:-6, 6:-6, 6:-6] = rng.standard_normal((nz - 12, ny - 12, nx - 12))
    u1 = u0 * 0.99
    st = [u1.copy(), u0.copy(), 0.5 * u1, 0.5 * u0]
    n = 60
    orc.steps3d_f64(st, c11, c13, c33, npad, 12.5, 12.5, 12.5, 1e-3, n, absorbing=False,
                    gradient_form=gradient_form)
    # after an even number of levels the arrays hold (level n+1, level n): swap roles to go back
    cx, px, cz, pz = st
    back = [px, cx, pz, cz]
    orc.steps3d_f64(back, c11, c13, c33, npad, 12.5, 12.5, 12.5, 1e-3, n, absorbing=False,
                    gradient_form=gradient_form)
    # n more levels (even) leave the latest level (level 0) in the first array
    assert rel(back[0], u0) < 1e-9 and rel(back[1], u1) < 1e-9
    assert rel(back[2], 0.5 * u0) < 1e-9


def test_modes_agree_short_run(orc):
    nz, ny, nx, npad = 40, 36, 44, 8
    v, e, dl = layered_model(nz, ny, nx, seed=6)
    c11, c13, c33 = cij_from(v, e, dl)
    nt = 130
    w = ricker(nt, 1e-3, 15.0, 0.08)
    sou = 20 + 18 * nx + 13 * nx * ny
    rec = np.array([ix + 18 * nx + 9 * nx * ny for ix in range(4, nx - 4)], np.int64)
    lmax = orc.lmax3d(nt)
    args = (c11, c13, c33, npad, 12.5, 12.5, 12.5, 1e-3, lmax, w, sou, rec, 4, (nt - 1) // 4 + 1)
    t32, _, _ = orc.forward3d(orc.F32, *args)
    tref, _, _ = orc.forward3d(orc.REF, *args)
    _, t64, _ = orc.forward3d(orc.F64, *args)
    assert np.abs(t64).max() > 0
    assert rel(tref, t64) < 1e-5
    assert rel(t32, t64) < 2e-5


def test_rtm_linear_in_data(orc):
    from common3d_drivers import problem

    P = problem(nt=130)
    rec3 = np.ascontiguousarray(P["recloc"][:, :3])
    nr = len(rec3)
    nnt = (P["nt"] - 1) // 4 + 1
    rng = np.random.default_rng(9)
    o1 = rng.standard_normal(nr * nnt).astype(np.float32) * 1e-9
    o2 = rng.standard_normal(nr * nnt).astype(np.float32) * 1e-9
    base = (P["souloc"], np.array([0]), rec3, P["w"], *P["cij"], P["nt"], P["npad"], P["o"], P["d"], P["dt"],
            P["rate"])
    i1 = orc.rtm3d(orc.F64, *base, o1)
    i2 = orc.rtm3d(orc.F64, *base, o2)
    i12 = orc.rtm3d(orc.F64, *base, (2 * o1.astype(np.float64) - 3 * o2).astype(np.float32))
    assert rel(i12, 2 * i1 - 3 * i2) < 1e-6
