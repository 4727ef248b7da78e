"""FP64 L-BFGS (vti_lbfgs_*) against the reference's lbfgs (lib/lbfgs).

tests/golden/reference_host.json holds the reference's own run (built from
lib/lbfgs/lbfgs_c.cpp, lsearch.cpp, la.cpp without OpenMP) on the test
functions of lib/lbfgs/functions.cpp, driven as test1c.cpp drives them: the
objective at every call, the number of calls, the final iflag and x.  The
functions are restated below (test infrastructure).  FP64 state makes the
iterates drift from the reference's FP32 ones, so the early objectives are
compared to 1e-4 and the end points to the textbook minimisers.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_host.json")))["lbfgs"]


# ---- lib/lbfgs/functions.cpp restated (double where the C++ promotes, float results)
def quad(x):
    x0, x1 = float(x[0]), float(x[1])
    g = np.array([2 * x0 - x1 - 4, 2 * x1 - x0 - 7], np.float32)
    return np.float32(x0 * x0 + x1 * x1 - x0 * x1 - 4 * x0 - 7 * x1), g


def camel6(x):
    x0, x1 = float(x[0]), float(x[1])
    x02, x12 = x0 * x0, x1 * x1
    x04 = x02 * x02
    f = (4.0 - 2.1 * x02 + x04 / 3.0) * x02 + x0 * x1 + 4 * (x12 - 1) * x12
    g = np.array([8 * x0 - 8.2 * x0 * x02 + 2 * x0 * x04 + x1, x0 + 16 * x1 * x12 - 8 * x1], np.float32)
    return np.float32(f), g


def rosenbrock(x):
    x = x.astype(np.float64)
    t1 = 1.0 - x[0::2]
    t2 = 10.0 * (x[1::2] - x[0::2] ** 2)
    g = np.empty_like(x)
    g[1::2] = 20.0 * t2
    g[0::2] = -2 * (x[0::2] * g[1::2] + t1)
    return np.float32(np.sum(t1 * t1 + t2 * t2)), g.astype(np.float32)


def powell(x):
    x = x.astype(np.float64)
    a, b, c, d = x[0::4], x[1::4], x[2::4], x[3::4]
    t1, t2, t3, t4 = a + 10 * b, c - d, b - 2 * c, a - d
    f = np.sum(t1 * t1 + 5 * t2 * t2 + t3 ** 4 + 10 * t4 ** 4)
    g = np.empty_like(x)
    g[0::4] = 2 * t1 + 40 * t4 ** 3
    g[1::4] = 20 * t1 + 4 * t3 ** 3
    g[2::4] = 10 * t2 - 8 * t3 ** 3
    g[3::4] = -10 * t2 - 40 * t4 ** 3
    return np.float32(f), g.astype(np.float32)


def trid(x):
    x = x.astype(np.float64)
    f = np.sum((x - 1) ** 2) - np.sum(x[:-1] * x[1:])
    g = 2 * (x - 1)
    g[:-1] -= x[1:]
    g[1:] -= x[:-1]
    return np.float32(f), g.astype(np.float32)


def rosenbrock1(x):
    x = x.astype(np.float64)
    t1, t2 = x[:-1] - 1, x[:-1] ** 2 - x[1:]
    g = np.zeros_like(x)
    g[:-1] += 2 * t1 + 400 * t2 * x[:-1]
    g[1:] -= 200 * t2
    return np.float32(np.sum(t1 * t1 + 100 * t2 * t2)), g.astype(np.float32)


def start(func, n):
    x = np.zeros(n, np.float32)
    if func in (0, 1):
        x[:] = (-1.0, -0.5)
    elif func == 2:
        x[:] = 2.0
    elif func == 3:
        x[0::4], x[1::4], x[3::4] = 3.0, -1.0, 1.0
    return x


FUNCS = {0: quad, 1: camel6, 2: rosenbrock, 3: powell, 4: trid, 5: rosenbrock1}
MINIMA = {
    0: lambda n: np.array([5.0, 6.0]),
    2: lambda n: np.ones(n),
    3: lambda n: np.zeros(n),
    4: lambda n: np.array([(i + 1) * (n - i) for i in range(n)], np.float64),  # x_i = i(n+1-i)
    5: lambda n: np.ones(n),
}


def run(vti, func, n, unit_step=False, maxcall=201):
    x = start(func, n)
    opt = vti.LBFGS(n, 5, unit_step=unit_step)
    fs, iflag, err = [], 0, None
    while True:
        f, g = FUNCS[func](x)
        fs.append(float(f))
        try:
            iflag = opt.step(x, float(f), g, iflag)
        except vti.VtiError as e:
            iflag, err = -1, str(e)
        if iflag != 1 or len(fs) >= maxcall:
            break
    return x, fs, iflag, err


def finished(iflag, err):
    """Converged, or stopped by a line-search exit that rounding forces (info 6,
    or info 4: the step shrank to its lower bound): with a float objective the
    last decreases fall below its resolution, which ends the reference's own
    quad and trid runs the same way.  The callers check where x ended."""
    return iflag == 0 or (iflag == -1 and ("info=6" in (err or "") or "info=4" in (err or "")))


@pytest.mark.parametrize("func", [0, 1, 2, 3, 4, 5])
def test_lbfgs_follows_reference(vti, func):
    ref = GOLD[str(func)]
    x, fs, iflag, err = run(vti, func, ref["n"])
    # the same iterates while FP32 and FP64 rounding have not yet parted them
    k = min(8, len(fs), len(ref["f"]))
    np.testing.assert_allclose(fs[:k], ref["f"][:k], rtol=5e-4, atol=1e-6)
    if ref["calls"] < 201:
        assert finished(iflag, err), (len(fs), iflag, err)
    xr = np.array(ref["x"])
    if func in MINIMA:
        xm = MINIMA[func](ref["n"])
        scale = max(1.0, np.abs(xm).max())
        tol = 5e-3 if func == 3 else 2e-3  # Powell's minimiser is singular: flat to fourth order
        assert np.abs(x - xm).max() <= np.abs(xr - xm).max() + tol * scale, (np.abs(x - xm).max(),
                                                                             np.abs(xr - xm).max())
    else:
        assert np.abs(x - xr).max() < 1e-3


@pytest.mark.parametrize("func", [0, 2, 4, 5])
def test_lbfgs_unit_step_reaches_textbook_minima(vti, func):
    n = GOLD[str(func)]["n"]
    x, fs, iflag, err = run(vti, func, n, unit_step=True, maxcall=2000)
    assert finished(iflag, err), (len(fs), iflag, err)
    xm = MINIMA[func](n)
    assert np.abs(x - xm).max() < 1e-3 * max(1.0, np.abs(xm).max())
    if GOLD[str(func)]["iflag"] == 0:
        # Liu-Nocedal restarts at stp = 1 take fewer calls than the reference's carried-over step
        assert len(fs) < GOLD[str(func)]["calls"]


def test_lbfgs_zero_gradient_is_done(vti):
    x = np.array([5.0, 6.0], np.float32)
    opt = vti.LBFGS(2, 5)
    f, g = quad(x)
    assert not g.any()
    assert opt.step(x, float(f), g, 0) == 0
    assert np.array_equal(x, [5.0, 6.0])


def test_lbfgs_rejects_bad_iflag(vti):
    x = start(0, 2)
    opt = vti.LBFGS(2, 5)
    f, g = quad(x)
    with pytest.raises(vti.VtiError):
        opt.step(x, float(f), g, 2)
    with pytest.raises(vti.VtiError):
        opt.step(x, float(f), g, 1)  # re-entry before a start
