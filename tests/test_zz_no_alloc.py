"""vti2d_synthetic_noalloc: the allocation-free form of synthetic_f.

Each case calls the entry point twice in a row through the C ABI -- a
larger input, then a smaller one, so the second call runs over what the
first left in the shared scratch array -- and checks both against
vti2d_synthetic on the same input, bit for bit (the one-shot schedule of
both is the same arithmetic).  The oracle's FP32 mode pins the first call as
well.  Sizes above the scratch limits are rejected with an error, before any
work is queued."""
import numpy as np
import pytest

from common2d import args2d, problem2d

pytestmark = pytest.mark.gpu


def _say(*a):
    print("case", *a, flush=True)


def _run(vti, ctx, P, noalloc, **kw):
    return vti.synthetic2d(ctx, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P), noalloc=noalloc, **kw)


@pytest.mark.parametrize("rate, ot", [(4e-3, 0.0), (1e-3, 0.0), (4e-3, -0.02)], ids=["rate4", "rate1", "ot<0"])
def test_two_calls_equal_allocating_entry(vti, orc, ctx, rate, ot):
    big = problem2d(vti, nx=120, nz=96, npad=16, nt=301, ns=3, seed=11, rate=rate)
    small = problem2d(vti, nx=48, nz=40, npad=8, nt=161, ns=2, seed=12, rate=rate)
    _say("noalloc big", big["nnz"], big["nnx"], big["nt"], "rate", rate, "ot", ot)
    a1 = _run(vti, ctx, big, True, ot=ot)
    b1 = _run(vti, ctx, big, False, ot=ot)
    o1 = orc.synthetic2d(orc.F32, *big["c"], big["w"], big["sloc"], big["rloc"], big["taper"], *args2d(big), ot=ot)
    assert np.abs(o1).max() > 0
    assert np.array_equal(a1, b1)
    assert np.array_equal(a1, o1)
    _say("noalloc small", small["nnz"], small["nnx"], small["nt"])
    a2 = _run(vti, ctx, small, True, ot=ot)
    b2 = _run(vti, ctx, small, False, ot=ot)
    assert np.abs(b2).max() > 0
    assert np.array_equal(a2, b2)


def test_rows_never_recorded_keep_their_values(vti, ctx):
    P = problem2d(vti, nx=40, nz=32, npad=8, nt=121, ns=2, seed=4)
    _say("noalloc keep rows", P["nnz"], P["nnx"])
    ratio = int(P["rate"] / P["dt"] + 0.5)
    shape = ((P["nt"] - 1) // ratio + 1, P["nr"])
    fill = np.full(shape, 7.25, np.float32)
    a = _run(vti, ctx, P, True, data=fill.copy())
    b = _run(vti, ctx, P, False, data=fill.copy())
    assert np.array_equal(a, b)
    assert np.all(a[0] == 7.25)  # row 0 is never recorded


def test_limits_rejected(vti, ctx):
    # 1040 x 1040 padded points > 2^20: rejected before any device work
    P = problem2d(vti, nx=16, nz=16, npad=8, nt=21, ns=1, seed=2)
    P = dict(P, nx=1024, nz=1024)
    with pytest.raises(vti.VtiError, match="scratch limit"):
        vti.synthetic2d(ctx, *(np.zeros((1040, 1040), np.float32) for _ in range(3)), P["w"], P["sloc"], P["rloc"],
                        np.ones((1040, 1040), np.float32), *args2d(P), noalloc=True)
