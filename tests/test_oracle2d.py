"""Pins of the 2D oracle (lib/vti-fwi-2d restated) that need no GPU.

- forwardCD (kernels.cu:452-498) and backwardDC (kernels.cu:104-148) against
  the discrete plane-wave solution of the coupled p/q system, with the
  8th-order weights as exact fractions (-205/72, 8/5, -1/5, 8/315, -1/560);
- backwardDC is the exact transpose of forwardCD (dot-product test);
- the objective of objFuncGradient_f equals 0.5*|synthetic_f - data|^2;
- Taylor test of the gradient and the Gauss-Newton product (FP64 oracle);
- FP32 oracle (the reference's arithmetic) within tolerance of FP64.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel

W = [Fraction(-205, 72), Fraction(8, 5), Fraction(-1, 5), Fraction(8, 315), Fraction(-1, 560)]


def _symbol(k, h):
    """-(second-derivative stencil) applied to cos(k x): a >= 0."""
    s = float(W[0]) + 2 * sum(float(W[j]) * np.cos(j * k * h) for j in range(1, 5))
    return -s / (h * h)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _step(orc, which, cur, prev, c, dx, dz, dt):
    L = orc._lib2()
    nnz, nnx = cur[0].shape
    nxt = [prev[0].copy(), prev[1].copy()]
    cc = [np.ascontiguousarray(x, np.float64) for x in c]
    fn = L.orc2d_forwardCD_f64 if which == "fwd" else L.orc2d_backwardDC_f64
    fn(_p(nxt[0]), _p(nxt[1]), _p(cur[0]), _p(cur[1]), *[_p(x) for x in cc], dx * dx, dz * dz, dt * dt, nnx, nnz)
    return nxt


@pytest.mark.parametrize("which", ["fwd", "bwd"])
def test_plane_wave_step(orc, which):
    nnz, nnx, dx, dz, dt = 40, 44, 10.0, 12.0, 1e-3
    v, eps, dl = 2200.0, 0.2, 0.08
    c33 = v * v
    c11, c13 = c33 * (1 + 2 * eps), c33 * np.sqrt(1 + 2 * dl)
    kx, kz = 0.7 / dx, 0.45 / dz
    a, b = _symbol(kx, dx), _symbol(kz, dz)
    # p_tt = -(M u) for u = (P, Q) e^{i(kx x + kz z - w t)}: forwardCD applies C.D, backwardDC D.C = (C.D)^T
    M = np.array([[c11 * a, c13 * b], [c13 * a, c33 * b]])
    if which == "bwd":
        M = M.T
    lam, vec = np.linalg.eig(M)
    j = int(np.argmax(lam.real))
    lam, vec = lam[j].real, vec[:, j].real
    w = np.arccos(1 - dt * dt * lam / 2) / dt
    X, Z = np.meshgrid(np.arange(nnx) * dx, np.arange(nnz) * dz)

    def field(t):
        ph = np.cos(kx * X + kz * Z - w * t)
        return [vec[0] * ph, vec[1] * ph]

    cs = [np.full((nnz, nnx), x) for x in (c11, c13, c33)]
    nxt = _step(orc, which, field(0.0), field(-dt), cs, dx, dz, dt)
    exp = field(dt)
    sl = (slice(4, nnz - 4), slice(4, nnx - 4))
    for k in range(2):
        err = np.abs(nxt[k][sl] - exp[k][sl]).max()
        assert err < 1e-9 * np.abs(vec).max(), (which, k, err)
    # the transposed system's plane wave is not a solution (c13 sits where it should)
    lam2, vec2 = np.linalg.eig(M.T)
    j2 = int(np.argmax(lam2.real))
    vec2 = vec2[:, j2].real
    w2 = np.arccos(1 - dt * dt * lam2[j2].real / 2) / dt

    def field2(t):
        ph = np.cos(kx * X + kz * Z - w2 * t)
        return [vec2[0] * ph, vec2[1] * ph]

    nxt2 = _step(orc, which, field2(0.0), field2(-dt), cs, dx, dz, dt)
    exp2 = field2(dt)
    assert max(np.abs(nxt2[k][sl] - exp2[k][sl]).max() for k in range(2)) > 1e-4 * np.abs(vec2).max()


def test_backwardDC_is_transpose_of_forwardCD(orc):
    rng = np.random.default_rng(5)
    nnz, nnx, dx, dz, dt = 36, 40, 10.0, 8.0, 1e-3
    c = [rng.uniform(2e6, 9e6, (nnz, nnx)) for _ in range(3)]
    c[1] = np.minimum(c[1], np.sqrt(c[0] * c[2]) * 0.9)
    sup = np.zeros((nnz, nnx))
    sup[8:-8, 8:-8] = 1
    x = [rng.standard_normal((nnz, nnx)) * sup for _ in range(2)]
    y = [rng.standard_normal((nnz, nnx)) * sup for _ in range(2)]
    zero = [np.zeros((nnz, nnx)), np.zeros((nnz, nnx))]
    Ax = _step(orc, "fwd", x, zero, c, dx, dz, dt)
    Aty = _step(orc, "bwd", y, zero, c, dx, dz, dt)
    lhs = sum(np.vdot(Ax[k] - 2 * x[k], y[k]) for k in range(2))
    rhs = sum(np.vdot(x[k], Aty[k] - 2 * y[k]) for k in range(2))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    # the pair is not self-adjoint: forwardCD is not its own transpose here
    Ay = _step(orc, "fwd", y, zero, c, dx, dz, dt)
    rhs_wrong = sum(np.vdot(x[k], Ay[k] - 2 * y[k]) for k in range(2))
    assert abs(lhs - rhs_wrong) > 1e-6 * abs(lhs)


@pytest.fixture(scope="module")
def P(vti):
    return problem2d(vti)


@pytest.fixture(scope="module")
def obs(orc, P):
    return orc.synthetic2d(orc.F64, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)).astype(
        np.float32)


def test_synthetic_f32_vs_f64(orc, P):
    d32 = orc.synthetic2d(orc.F32, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    d64 = orc.synthetic2d(orc.F64, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    assert np.abs(d64).max() > 0
    assert not d64[0].any()  # sample 0 is never recorded for ratio > 1 (synthetic_f.cu:97)
    # the reference's 2D path is FP32 throughout; storing the wavefield in FP32
    # every step leaves its traces 2.0e-5 (rel L2) from FP64 after 200 steps
    assert rel(d32, d64) < 4e-5


def test_synthetic_linear_in_wavelet(orc, P):
    d1 = orc.synthetic2d(orc.F64, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    # the source term dt^2*w is formed in FP32 (synthetic_f.cu:102): scale by a power of two
    d4 = orc.synthetic2d(orc.F64, *P["c"], 4 * P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    assert rel(d4, 4 * d1) < 1e-14


@pytest.mark.parametrize("mode", ["F64", "F32"])
def test_objective_matches_synthetic(orc, P, obs, mode):
    m = getattr(orc, mode)
    f, _ = orc.fg2d(m, False, obs, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    d = orc.synthetic2d(m, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    r = d.astype(np.float64) - obs
    ref = 0.5 * float(np.sum(r * r))
    assert abs(f - ref) <= (1e-12 if mode == "F64" else 1e-6) * ref, (f, ref)


def _smooth_bump(P, k, seed, sigma=8.0, amp=0.01):
    rng = np.random.default_rng(seed)
    nnz, nnx = P["nnz"], P["nnx"]
    Z, X = np.mgrid[0:nnz, 0:nnx]
    cz, cx = rng.uniform(0.45, 0.55) * nnz, rng.uniform(0.45, 0.55) * nnx
    b = np.exp(-((Z - cz) ** 2 + (X - cx) ** 2) / (2 * sigma ** 2))
    dc = [np.zeros((nnz, nnx)) for _ in range(3)]
    dc[k] = b * P["c"][k].mean() * amp
    return dc


@pytest.fixture(scope="module")
def P2(vti):
    """rate = 2 dt (the smallest snapshot ratio the reference allows)."""
    return problem2d(vti, rate=2e-3)


@pytest.fixture(scope="module")
def obs2(orc, P2):
    return orc.synthetic2d(orc.F64, *P2["ctrue"], P2["w"], P2["sloc"], P2["rloc"], P2["taper"], *args2d(P2)).astype(
        np.float32)


def _fd_born(orc, P, dc, h=1.0):
    """Central difference of synthetic_f; h=1 keeps the FP32 rounding of c + h*dc
    (the drivers take FP32 models) well below the O(h^2) term."""
    dp = orc.synthetic2d(orc.F64, *[P["c"][i] + h * dc[i] for i in range(3)], P["w"], P["sloc"], P["rloc"],
                         P["taper"], *args2d(P))
    dm = orc.synthetic2d(orc.F64, *[P["c"][i] - h * dc[i] for i in range(3)], P["w"], P["sloc"], P["rloc"],
                         P["taper"], *args2d(P))
    return (dp - dm) / (2 * h)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_gradient_taylor(orc, P2, obs2, k):
    """dJ/dc . dc against a central difference of J = 0.5|synthetic - data|^2
    (FP64 oracle).  objFuncGradient_f injects the residual at every time step,
    interpolated between samples (injectData, kernels.cu:208-217), so what it
    returns approximates ratio * grad J, ratio = rate/dt; it also images
    against interpolated snapshots.  Measured at ratio 2: 0.95 ratio grad J
    for all three coefficients; the bar is 10%.  A dropped term, a sign or a
    swapped coefficient fails it."""
    P, obs = P2, obs2
    ratio = int(P["rate"] / P["dt"] + 0.5)
    _, g = orc.fg2d(orc.F64, False, obs, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    dc = _smooth_bump(P, k, 20 + k)
    d0 = orc.synthetic2d(orc.F64, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    fd = float(np.sum((d0 - obs) * _fd_born(orc, P, dc)))  # <r, J dc> = dJ/dh
    an = float(np.sum(g[k] * dc[k]))
    assert abs(an / (ratio * fd) - 1) < 0.1, (an, ratio * fd)


def test_gn_hessian_linear(orc, P):
    dc1, dc2 = _smooth_bump(P, 0, 31), _smooth_bump(P, 2, 32)

    def H(dc):
        return orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))[1]

    h1, h2, h12 = H(dc1), H(dc2), H([dc1[i] + 2 * dc2[i] for i in range(3)])
    assert rel(h12, h1 + 2 * h2) < 1e-11


def test_gn_hessian_is_its_adjoint_of_born_data(orc, P):
    """GNhessian_f = (the Hessian's forwardCD adjoint, hessian_f.cu:854) applied
    to the Born data of dc.  The Born data come independently from a central
    difference of synthetic_f; the adjoint runs through the gradient path with
    forwardCD (oracle check mode born=2) on residual = Born data."""
    dc = [_smooth_bump(P, k, 40 + k)[k] for k in range(3)]
    ns = len(P["sloc"])
    Hdc = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P),
                   ngpu=ns)[1]
    d0 = orc.synthetic2d(orc.F64, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    born = _fd_born(orc, P, dc)
    # residual d0 - data = born; data rows 0 stay as d0 (both zero there)
    _, A = orc.fg2d(orc.F64, 2, (d0 - born).astype(np.float32), *P["c"], None, P["w"], P["sloc"], P["rloc"],
                    P["taper"], *args2d(P))
    assert np.abs(Hdc).max() > 0
    assert rel(Hdc, A) < 5e-4, rel(Hdc, A)  # measured 1.0e-4
    # it is not the gradient path's exact adjoint (backwardDC) applied to the same data
    _, G = orc.fg2d(orc.F64, 0, (d0 - born).astype(np.float32), *P["c"], None, P["w"], P["sloc"], P["rloc"],
                    P["taper"], *args2d(P))
    assert rel(Hdc, G) > 1e-2


def test_gn_hessian_carries_born_state_between_shots_on_one_gpu(orc, P):
    """GNhessian_f never refreshes d_Ddpq (D of the Born field) when a GPU
    starts its next shot (hessian_f.cu:674-675,757,770): the first Born step
    of that shot reads the previous shot's last D.  With one shot per GPU the
    product is the adjoint of the Born data (above); with two shots on one GPU
    it is not.  The HIP path computes every shot from a clean state, i.e. the
    reference as run with one GPU per shot."""
    dc = [_smooth_bump(P, k, 40 + k)[k] for k in range(3)]
    ns = len(P["sloc"])
    h1 = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P), ngpu=1)[1]
    hn = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P),
                  ngpu=ns)[1]
    assert rel(h1, hn) > 1e-2
    # the first shot is the same either way
    one = P["sloc"][:1]
    a = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], one, P["rloc"], P["taper"], *args2d(P), ngpu=1)[1]
    b = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], one, P["rloc"], P["taper"], *args2d(P), ngpu=ns)[1]
    assert np.array_equal(a, b)


def test_gradient_does_not_depend_on_gpu_count(orc, P, obs):
    """objFuncGradient_f carries no state between shots that changes results:
    one GPU or one GPU per shot differ only by FP32 summation order."""
    ns = len(P["sloc"])
    f1, g1 = orc.fg2d(orc.F32, False, obs, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    fn, gn = orc.fg2d(orc.F32, False, obs, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P),
                      ngpu=ns)
    assert f1 == fn
    assert rel(g1, gn) < 1e-6


def test_fg_f32_vs_f64(orc, P, obs):
    """The reference's FP32 arithmetic against FP64.  With zero data the
    residual is the modelled data and the gradient sits 1.1e-5 from FP64.
    With data from a model 5% off in a blob the residual is 0.2% of the data,
    so the FP32 trace error (2e-5 of the data) is 1% of the residual, and the
    gradient follows it (measured 1.05e-2)."""
    zero = np.zeros_like(obs)
    f32, g32 = orc.fg2d(orc.F32, False, zero, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    f64, g64 = orc.fg2d(orc.F64, False, zero, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    assert abs(f32 - f64) < 1e-5 * f64
    assert rel(g32, g64) < 1e-4
    f32, g32 = orc.fg2d(orc.F32, False, obs, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    f64, g64 = orc.fg2d(orc.F64, False, obs, *P["c"], None, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    assert abs(f32 - f64) < 2e-2 * f64
    assert rel(g32, g64) < 3e-2


def test_gn_f32_vs_f64(orc, P):
    rng = np.random.default_rng(1)
    dc = [rng.standard_normal((P["nnz"], P["nnx"])) * P["c"][k].mean() * 0.01 for k in range(3)]
    ns = len(P["sloc"])
    h32 = orc.fg2d(orc.F32, True, None, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P), ngpu=ns)[1]
    h64 = orc.fg2d(orc.F64, True, None, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P), ngpu=ns)[1]
    assert rel(h32, h64) < 1e-4  # measured 3.6e-6
