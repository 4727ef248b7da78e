"""The oracle's restatement of rtm_f / odcig_f (lib/vti-fwi-2d/rtm_f.cu,
odcig_f.cu), pinned by relations the reference's algorithm fixes:
- the image is linear in the data, and zero for zero data;
- with wavelet[0] = 0, odcig_f's zero-offset plane is rtm_f's image / dt^2
  (odcig_f injects wavelet[it-1] where rtm_f injects dt^2*wavelet[it-1]);
- data recorded from ot = -k*(rate/dt)*dt, with the first k rows cut, image
  as data from ot = 0 whose first k rows are zero (the ntNeg row alignment);
- the image of a two-shot run equals the sum of the one-shot images.
All in the oracle's FP64 mode."""
import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d import rel


@pytest.fixture(scope="module")
def P(vti):
    return problem2d(vti, nx=48, nz=36, npad=10, nt=121, ns=2)


@pytest.fixture(scope="module")
def obs(orc, P):
    return orc.synthetic2d(orc.F64, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)).astype(
        np.float32)


def rtm(orc, P, data, **kw):
    return orc.rtm2d(orc.F64, data, *P["c"], kw.pop("w", P["w"]), kw.pop("sloc", P["sloc"]), P["rloc"], P["taper"],
                     *args2d(P), **kw)


def test_zero_data_zero_image(orc, P, obs):
    assert np.all(rtm(orc, P, np.zeros_like(obs)) == 0)


def test_linear_in_data(orc, P, obs):
    """Superposition: data on the even receivers plus data (scaled by 2, exact
    in float) on the odd ones."""
    rng = np.random.default_rng(1)
    d1 = obs.copy()
    d1[:, 1::2] = 0.0
    d2 = rng.standard_normal(obs.shape).astype(np.float32)
    d2[:, 0::2] = 0.0
    a = rtm(orc, P, d1)
    b = rtm(orc, P, d2)
    ab = rtm(orc, P, d1 + np.float32(2) * d2)
    assert np.abs(a).max() > 0 and np.abs(b).max() > 0
    assert rel(ab, a + 2 * b) < 1e-12


def test_odcig_zero_offset_is_rtm_over_dt2(orc, P, obs):
    w = P["w"].copy()
    w[0] = 0.0
    img = rtm(orc, P, obs, w=w)
    ocg = rtm(orc, P, obs, w=w, nhx=3)
    dt2 = np.float32(P["dt"]) * np.float32(P["dt"])
    # the sources agree up to the float rounding of dt2*wavelet
    assert rel(ocg[3] * dt2, img) < 1e-6
    assert np.abs(ocg[0]).max() > 0 and np.abs(ocg[6]).max() > 0


def test_negative_start_time_alignment(orc, P, obs):
    k = 3
    ratio = int(P["rate"] / P["dt"] + 0.5)
    a = list(args2d(P))
    cut = obs[k:]
    img_neg = orc.rtm2d(orc.F64, cut, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *a, ot=-k * ratio * P["dt"])
    zeroed = obs.copy()
    zeroed[:k] = 0.0
    img0 = rtm(orc, P, zeroed)
    assert np.abs(img0).max() > 0
    # the interpolation weights are formed in float from shifted times: ~1e-7
    assert rel(img_neg, img0) < 1e-5


def test_shots_add(orc, P, obs):
    both = rtm(orc, P, obs)
    one = rtm(orc, P, obs, sloc=P["sloc"][:1])
    two = rtm(orc, P, obs, sloc=P["sloc"][1:])
    assert rel(both, one + two) < 1e-13
