"""The oracle's restatement of hessian_f (lib/vti-fwi-2d/hessian_f.cu:15-573),
the full Hessian-vector product, pinned in FP64:
- with its Born snapshot pipeline seeded from the Born snapshots, it is
  linear in the perturbation dc and it is the derivative of the oracle's
  objFuncGradient_f gradient along dc (central differences);
- the reference seeds that buffer from the background snapshots instead
  (hessian_f.cu:381-382): in the imaging of the residual's adjoint field one
  Born snapshot slot is replaced by a background one, so the product becomes
  affine in dc -- nonzero for dc = 0, where the seeded variant gives 0."""
import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel


@pytest.fixture(scope="module")
def P(vti):
    return problem2d(vti, nx=40, nz=32, npad=10, nt=121, ns=1)


@pytest.fixture(scope="module")
def obs(orc, P):
    return orc.synthetic2d(orc.F64, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)).astype(
        np.float32)


def dc_of(P, seed, mask=None):
    rng = np.random.default_rng(seed)
    dc = [(0.05 * c * rng.standard_normal(c.shape)).astype(np.float32) for c in P["c"]]
    if mask is not None:
        dc = [np.where(mask, d, 0).astype(np.float32) for d in dc]
    return dc


def H(orc, P, obs, dc, seed_bg=True):
    return orc.hessian2d(orc.F64, obs, *P["c"], dc, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P),
                         seed_bg=seed_bg)


def test_linear_in_dc(orc, P, obs):
    even = (np.arange(P["nnx"])[None, :] % 2 == 0) & np.ones((P["nnz"], 1), bool)
    d1 = dc_of(P, 1, even)
    d2 = dc_of(P, 2, ~even)
    h1, h2 = H(orc, P, obs, d1, False), H(orc, P, obs, d2, False)
    h12 = H(orc, P, obs, [a + np.float32(2) * b for a, b in zip(d1, d2)], False)
    assert np.abs(h1).max() > 0 and np.abs(h2).max() > 0
    assert rel(h12, h1 + 2 * h2) < 1e-12


def test_derivative_of_the_gradient(orc, P, obs):
    dc = dc_of(P, 3)
    hv = H(orc, P, obs, dc, seed_bg=False)
    h = 0.02
    gp = orc.fg2d(orc.F64, False, obs, *[(c + h * d).astype(np.float32) for c, d in zip(P["c"], dc)], None, P["w"],
                  P["sloc"], P["rloc"], P["taper"], *args2d(P))[1]
    gm = orc.fg2d(orc.F64, False, obs, *[(c - h * d).astype(np.float32) for c, d in zip(P["c"], dc)], None, P["w"],
                  P["sloc"], P["rloc"], P["taper"], *args2d(P))[1]
    fd = (gp - gm) / (2 * h)
    e = rel(hv, fd)
    print(f"full Hessian vs central difference of the gradient: {e:.3e}")
    assert e < 5e-3


def test_background_seed_quirk_makes_it_affine(orc, P, obs):
    zero = [np.zeros_like(c) for c in P["c"]]
    h0 = H(orc, P, obs, zero)
    assert np.all(H(orc, P, obs, zero, seed_bg=False) == 0)
    assert np.abs(h0).max() > 0
    even = (np.arange(P["nnx"])[None, :] % 2 == 0) & np.ones((P["nnz"], 1), bool)
    d1 = dc_of(P, 1, even)
    d2 = dc_of(P, 2, ~even)
    h1, h2 = H(orc, P, obs, d1), H(orc, P, obs, d2)
    h12 = H(orc, P, obs, [a + np.float32(2) * b for a, b in zip(d1, d2)])
    assert rel(h12, h1 + 2 * h2 - 2 * h0) < 1e-12
