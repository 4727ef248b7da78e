"""Device-resident FP64 L-BFGS (vti_lbfgs_create_device) against the host
form of the same method (vti_lbfgs_create, itself checked against the
reference's lbfgs in test_lbfgs.py).

Both run the same line search on the same scalars; they differ only in the
order the dot products are summed (index order on the host, a fixed
block/tree order on the device).  So:
- on 2-variable problems (every sum has two terms, added the same way) the
  two runs are identical bit for bit;
- on the larger test functions the objective sequences agree to 1e-9 over
  the first iterations and both runs end at the textbook minimisers;
- x and g given as device pointers (hipMalloc) give the same bits as host
  arrays;
- on a 2-million-variable quadratic the device run repeats itself bit for
  bit (the fixed reduction order) and matches the host run to 1e-9.
"""
import ctypes

import numpy as np
import pytest

from test_lbfgs import FUNCS, GOLD, MINIMA, finished, start

pytestmark = pytest.mark.gpu


def _say(*a):
    print("case", *a, flush=True)


def run(vti, func, n, ctx=None, maxcall=201, unit_step=False):
    x = start(func, n)
    opt = vti.LBFGS(n, 5, unit_step=unit_step, ctx=ctx)
    fs, xs, iflag, err = [], [], 0, None
    while True:
        f, g = FUNCS[func](x)
        fs.append(float(f))
        try:
            iflag = opt.step(x, float(f), g, iflag)
        except vti.VtiError as e:
            iflag, err = -1, str(e)
        xs.append(x.copy())
        if iflag != 1 or len(fs) >= maxcall:
            break
    opt.close()
    return x, fs, xs, iflag, err


@pytest.mark.parametrize("func", [0, 1])
def test_two_variables_bit_identical(vti, ctx, func):
    n = GOLD[str(func)]["n"]
    _say("lbfgs device vs host, 2 variables, func", func)
    xh, fh, xsh, ih, eh = run(vti, func, n)
    xd, fd, xsd, idv, ed = run(vti, func, n, ctx=ctx)
    assert fh == fd and ih == idv
    for a, b in zip(xsh, xsd):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("func", [2, 3, 4, 5])
def test_larger_functions_follow_host(vti, ctx, func):
    n = GOLD[str(func)]["n"]
    _say("lbfgs device vs host, func", func, "n", n)
    xh, fh, _, ih, eh = run(vti, func, n, unit_step=True, maxcall=2000)
    xd, fd, _, idv, ed = run(vti, func, n, ctx=ctx, unit_step=True, maxcall=2000)
    k = min(12, len(fh), len(fd))
    np.testing.assert_allclose(fd[:k], fh[:k], rtol=1e-9, atol=1e-12)
    assert finished(idv, ed), (len(fd), idv, ed)
    xm = MINIMA[func](n)
    assert np.abs(xd - xm).max() < 1e-3 * max(1.0, np.abs(xm).max())


def _hip():
    """The HIP runtime (hipMalloc/hipMemcpy through ctypes): device buffers
    for x and g without a tensor library."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def test_device_pointers_equal_host_arrays(vti, ctx):
    """x and g in device memory (hipMalloc) give the iterates host arrays give."""
    func, n = 2, GOLD["2"]["n"]
    _say("lbfgs device pointers, rosenbrock n", n)
    _, fh, xsh, _, _ = run(vti, func, n, ctx=ctx, maxcall=40)
    hip = _hip()
    H2D, D2H = 1, 2
    xd, gd = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(xd), 4 * n) == 0 and hip.hipMalloc(ctypes.byref(gd), 4 * n) == 0
    try:
        x = start(func, n)
        assert hip.hipMemcpy(xd, x.ctypes.data, 4 * n, H2D) == 0
        opt = vti.LBFGS(n, 5, ctx=ctx)
        iflag, k = 0, 0
        while True:
            assert hip.hipMemcpy(x.ctypes.data, xd, 4 * n, D2H) == 0
            f, g = FUNCS[func](x)
            assert float(f) == fh[k]
            assert hip.hipMemcpy(gd, g.ctypes.data, 4 * n, H2D) == 0
            fl, it, nf = ctypes.c_int(iflag), ctypes.c_int(0), ctypes.c_int(0)
            rc = vti.lib.vti_lbfgs_step(opt.h, xd, float(f), gd, ctypes.byref(fl), ctypes.byref(it),
                                        ctypes.byref(nf))
            assert rc == 0, vti.last_error()
            iflag = fl.value
            assert hip.hipMemcpy(x.ctypes.data, xd, 4 * n, D2H) == 0
            assert np.array_equal(x, xsh[k])
            k += 1
            if iflag != 1 or k >= len(fh):
                break
        opt.close()
    finally:
        hip.hipFree(xd)
        hip.hipFree(gd)


def test_two_million_variables_reproducible(vti, ctx):
    """A separable quadratic with 2^21 variables (the size of a cfg2 model,
    3 x 592 x 1104): ten calls, twice on the device (bit-identical) and once
    on the host (to 1e-9)."""
    n = 1 << 21
    rng = np.random.default_rng(3)
    a = rng.uniform(0.5, 4.0, n).astype(np.float64)
    b = rng.standard_normal(n)

    def quad(x):
        xd = x.astype(np.float64)
        return float(0.5 * np.sum(a * xd * xd) - np.sum(b * xd)), (a * xd - b).astype(np.float32)

    def go(c):
        x = np.zeros(n, np.float32)
        opt = vti.LBFGS(n, 5, ctx=c)
        fs, iflag = [], 0
        for _ in range(10):
            f, g = quad(x)
            fs.append(f)
            iflag = opt.step(x, f, g, iflag)
            if iflag != 1:
                break
        opt.close()
        return fs, x

    _say("lbfgs 2^21 variables")
    f1, x1 = go(ctx)
    f2, x2 = go(ctx)
    fh, xh = go(None)
    assert f1 == f2 and np.array_equal(x1, x2)
    np.testing.assert_allclose(f1, fh, rtol=1e-9)
    assert f1[-1] < f1[0]
