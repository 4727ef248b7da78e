"""The AGC objective of the 3D drivers (lib/vti-fwi-3d/gpu/agc.cpp) in the
oracle: closed forms of the circular gain window and the adjoint source as
the exact gradient of half the AGC misfit (central differences, FP64)."""
import numpy as np
import pytest


def test_agc_constant_trace(orc):
    nt, h = 50, 4
    d = np.full((1, nt), 3.0)
    fi, _ = orc.agc_residual(orc.F64, d, np.zeros((1, nt), np.float32), h)
    # every window holds 2h+1 samples of 9: h(d) = 3/sqrt(9(2h+1)) everywhere
    assert abs(fi - nt / (2 * h + 1)) < 1e-12


@pytest.mark.parametrize("a,b,wrapped", [(46, 2, True), (45, 2, False), (10, 17, False), (20, 26, True)])
def test_agc_window_is_circular(orc, a, b, wrapped):
    """Spikes 1 at a and 2 at b: inside one window (circular distance <= h)
    h(d) = (1, 2)/sqrt(5) and the misfit is 1; apart, each is its own window
    and the misfit is 2."""
    nt, h = 50, 6
    d = np.zeros((1, nt))
    d[0, a], d[0, b] = 1.0, 2.0
    fi, _ = orc.agc_residual(orc.F64, d, np.zeros((1, nt), np.float32), h)
    assert abs(fi - (1.0 if wrapped else 2.0)) < 1e-12


def test_agc_adjoint_is_gradient(orc):
    rng = np.random.default_rng(3)
    ntr, nt, h = 3, 40, 5
    d = rng.standard_normal((ntr, nt))
    d0 = rng.standard_normal((ntr, nt)).astype(np.float32)
    fi, adj = orc.agc_residual(orc.F64, d, d0, h)
    eps = 1e-6
    for k in [(0, 3), (1, 0), (2, 39), (1, 20)]:
        dp, dm = d.copy(), d.copy()
        dp[k] += eps
        dm[k] -= eps
        fd = (orc.agc_residual(orc.F64, dp, d0, h)[0] - orc.agc_residual(orc.F64, dm, d0, h)[0]) / (4 * eps)
        assert abs(adj[k] - fd) < 1e-6 * max(1.0, abs(fd)), (k, adj[k], fd)


def test_agc_f32_close_to_f64(orc):
    rng = np.random.default_rng(4)
    d = rng.standard_normal((4, 60))
    d0 = rng.standard_normal((4, 60)).astype(np.float32)
    f32, a32 = orc.agc_residual(orc.F32, d.astype(np.float32), d0, 7)
    f64, a64 = orc.agc_residual(orc.F64, d.astype(np.float32).astype(np.float64), d0, 7)
    assert abs(f32 - f64) < 1e-5 * f64
    assert np.abs(a32 - a64).max() < 1e-5 * np.abs(a64).max()
