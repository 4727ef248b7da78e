"""The reference's CPU path restated (oracle/cpu3d.c, from
lib/vti-fwi-3d/cpu/laplacian3d.ispc and vtifd3d8o-modeling.cpp): the CPU
baseline of bench.py.  Pinned here by the FP64 oracle step (same equation,
FP64), by the v/eps/del form agreeing with the c_ij form, and by the closed
form of its damping profile."""
import numpy as np

from common3d import cij_from, layered_model, rel


def _fields(nz, ny, nx, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((nz, ny, nx)).astype(np.float32) * 1e-3 for _ in range(4)]


def test_cpu_path_matches_fp64_oracle(orc):
    """The CPU path updates z in [4, nz-4); the GPU path it is compared with
    (the FP64 oracle) updates every plane of its 4-plane blocks, so after k
    steps the two agree on the planes k*4 away from that difference."""
    nz, ny, nx = 40, 26, 30
    v, e, d = layered_model(nz, ny, nx, seed=4)
    c = cij_from(v, e, d)
    for nsteps in (1, 2):
        px, cx, pz, cz = _fields(nz, ny, nx, 1)
        st = [px.copy(), cx.copy(), pz.copy(), cz.copy()]
        orc.cpu3d_steps(1, st, *c, 0, 12.5, 12.5, 12.5, 1e-3, nsteps)
        s64 = [cx.astype(np.float64), px.astype(np.float64), cz.astype(np.float64), pz.astype(np.float64)]
        orc.steps3d_f64(s64, *c, 0, 12.5, 12.5, 12.5, 1e-3, nsteps, absorbing=False, gradient_form=True)
        new64 = (s64[0], s64[2]) if nsteps % 2 == 0 else (s64[1], s64[3])
        zs = slice(4 + 4 * nsteps, nz - 4 - 4 * nsteps)
        assert rel(st[1][zs], new64[0][zs]) < 1e-6
        assert rel(st[3][zs], new64[1][zs]) < 1e-6


def test_cpu_path_vepsdel_equals_cij(orc):
    nz, ny, nx = 20, 18, 22
    v, e, d = layered_model(nz, ny, nx, seed=5)
    c33 = (v * v).astype(np.float32)
    c11 = (c33 * (np.float32(1) + np.float32(2) * e)).astype(np.float32)
    c13 = (c33 * np.sqrt(np.float32(1) + np.float32(2) * d)).astype(np.float32)
    f = _fields(nz, ny, nx, 2)
    a = [x.copy() for x in f]
    b = [x.copy() for x in f]
    w = np.linspace(1, 2, 5).astype(np.float32)
    src = 11 + 9 * nx + 10 * nx * ny
    orc.cpu3d_steps(0, a, v, e, d, 5, 10.0, 10.0, 10.0, 1e-3, 5, src, w)
    orc.cpu3d_steps(1, b, c11, c13, c33, 5, 10.0, 10.0, 10.0, 1e-3, 5, src, w)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_cpu_damping_profile_closed_form(orc):
    npad = 30
    dmp = orc.cpu3d_damping(npad)
    i = np.arange(npad)
    exp = (0.95 + 0.05 * np.cos(3.14159265359 * (npad - i) / npad)).astype(np.float32)
    assert np.array_equal(dmp, exp)
    assert dmp[0] == np.float32(0.9) and np.all(np.diff(dmp) > 0)


def test_cpu_gradient_step_matches_fp64_formulas(orc):
    """One backward level of the CPU gradient (gradientCij + forward on the
    adjoint, no damping) against the same level written out in FP64 numpy:
    sigma' = 2 sigma + dt2 (c11 (dxx+dyy) + c13 dzz) - sigma_prev (the
    stencils divided by d2), lambda' by forward, and the imaging gc11 = lx
    (dxx+dyy), gc33 = lz dzz, gc13 = lx dzz + lz (dxx+dyy) at the updated
    interior."""
    import brute_history as bh

    nz, ny, nx = 20, 18, 22
    v, e, d = layered_model(nz, ny, nx, seed=7)
    c33 = v.astype(np.float64) ** 2
    c11 = c33 * (1 + 2 * e.astype(np.float64))
    c13 = c33 * np.sqrt(1 + 2 * d.astype(np.float64))
    sp, sc, zp, zc = _fields(nz, ny, nx, 3)
    lp, lc, mp, mc = _fields(nz, ny, nx, 4)
    sig, lam = [sp.copy(), sc.copy(), zp.copy(), zc.copy()], [lp.copy(), lc.copy(), mp.copy(), mc.copy()]
    gc = [np.zeros((nz, ny, nx), np.float32) for _ in range(3)]
    dx, dt = 10.0, 1e-3
    orc.cpu3d_backward_steps(sig, lam, gc, v, e, d, 0, dx, dx, dx, dt, 1)
    f64 = lambda a: a.astype(np.float64)  # noqa: E731
    h = dx * dx

    def terms(ux, uz):
        return (bh._d2(f64(ux), 2) + bh._d2(f64(ux), 1)) / h, bh._d2(f64(uz), 0) / h

    txy, tz = terms(sc, zc)
    dt2 = dt * dt
    want_sx = 2 * f64(sc) + dt2 * (c11 * txy + c13 * tz) - f64(sp)
    lxy, lz = terms(lc, mc)
    want_lx = 2 * f64(lc) + dt2 * (c11 * lxy + c13 * lz) - f64(lp)
    inner = (slice(4, nz - 4), slice(4, ny - 4), slice(4, nx - 4))
    # after one level the newest levels sit in the arrays passed as prev
    assert rel(sig[0][inner], want_sx[inner]) < 1e-5
    assert rel(lam[0][inner], want_lx[inner]) < 1e-5
    assert rel(gc[0][inner], (f64(lc) * txy)[inner]) < 1e-5
    assert rel(gc[2][inner], (f64(mc) * tz)[inner]) < 1e-5
    assert rel(gc[1][inner], (f64(lc) * tz + f64(mc) * txy)[inner]) < 1e-5
