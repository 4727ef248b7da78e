"""2D RTM and ODCIG (rtm_f, odcig_f) on the GPU against the oracle's FP32
restatement, bit for bit with one shot per GPU in the oracle (the HIP path
sums shots in FP64, exact for two), with each result's error against the
FP64 oracle stated."""
import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def P(vti):
    return problem2d(vti, nx=64, nz=48, npad=12, nt=201, ns=2)


def _obs(orc, P, rate, ot):
    a = list(args2d(P))
    a[-1] = rate
    return orc.synthetic2d(orc.F64, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *a, ot=ot).astype(
        np.float32), a


@pytest.mark.parametrize("nhx", [-1, 3, 6], ids=["rtm", "odcig3", "odcig6"])
@pytest.mark.parametrize("rate,ot", [(4e-3, 0.0), (4e-3, -0.02), (1e-3, 0.0)])
@pytest.mark.parametrize("nshot", [1, 2])
def test_rtm_vs_oracle(vti, orc, ctx, P, nshot, rate, ot, nhx):
    _say("2d rtm nshot", nshot, "rate", rate, "ot", ot, "nhx", nhx)
    data, a = _obs(orc, P, rate, ot)
    sloc = P["sloc"][:nshot]
    img = vti.rtm2d(ctx, data, *P["c"], P["w"], sloc, P["rloc"], P["taper"], *a, ot=ot, nhx=nhx)
    o32 = orc.rtm2d(orc.F32, data, *P["c"], P["w"], sloc, P["rloc"], P["taper"], *a, ot=ot, nhx=nhx, ngpu=nshot)
    o64 = orc.rtm2d(orc.F64, data, *P["c"], P["w"], sloc, P["rloc"], P["taper"], *a, ot=ot, nhx=nhx)
    assert np.abs(o32).max() > 0
    assert np.array_equal(img, o32)
    e = rel(img, o64)
    _say(f"  rel vs FP64 {e:.3e}")
    assert e < 1e-4


def test_rtm_rejects_stacked_receivers(vti, ctx, P):
    rl = P["rloc"].copy()
    rl[1] = (rl[0][0], rl[0][1] + 1)  # right below receiver 0
    with pytest.raises(vti.VtiError, match="below"):
        vti.rtm2d(ctx, np.zeros(10**5, np.float32), *P["c"], P["w"], P["sloc"][:1], rl, P["taper"], *args2d(P))
