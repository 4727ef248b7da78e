"""The 2D full Hessian-vector product (hessian_f) on the GPU against the
oracle's FP32 restatement -- bit for bit with one shot per GPU in the oracle
(the HIP path sums shots in FP64, exact for two) -- and its error against the
FP64 oracle."""
import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P(vti):
    return problem2d(vti, nx=64, nz=48, npad=12, nt=201, ns=2)


@pytest.mark.parametrize("nshot", [1, 2])
def test_full_hessian_vs_oracle(vti, orc, ctx, P, nshot):
    print("2d full Hessian nshot", nshot, flush=True)
    obs = orc.synthetic2d(orc.F64, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)).astype(
        np.float32)
    rng = np.random.default_rng(7)
    dc = [(0.05 * c * rng.standard_normal(c.shape)).astype(np.float32) for c in P["c"]]
    sloc = P["sloc"][:nshot]
    h = vti.hessian2d(ctx, obs, *P["c"], *dc, P["w"], sloc, P["rloc"], P["taper"], *args2d(P))
    o32 = orc.hessian2d(orc.F32, obs, *P["c"], dc, P["w"], sloc, P["rloc"], P["taper"], *args2d(P), ngpu=nshot)
    o64 = orc.hessian2d(orc.F64, obs, *P["c"], dc, P["w"], sloc, P["rloc"], P["taper"], *args2d(P))
    assert np.abs(o32).max() > 0
    assert np.array_equal(h, o32)
    e = rel(h, o64)
    print(f"  rel vs FP64 {e:.3e}", flush=True)
    assert e < 2e-4  # measured 6.5e-5 (1 shot), 4.7e-5 (2 shots)
