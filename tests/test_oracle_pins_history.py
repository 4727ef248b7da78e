"""Pins of the oracle's 3D gradient / RTM imaging, receiver weights, residual
sign, t^p gain and time alignment, and of its 2D RTM / ODCIG imaging (CPU).

Each test runs a tiny problem through the oracle's FP64 mode and through
tests/brute_history.py, a numpy restatement that keeps every level of the
source wavefield instead of rebuilding it (random-boundary reversal in 3D,
slice interpolation in 2D), and requires agreement to roundoff.  The same
checks run against deliberately broken builds of the oracle in
test_oracle_mutants.py, where each must fail.
"""
import numpy as np
import pytest

import pins


def test_gradient3d_stored_history(orc):
    err, J_err = pins.gradient3d_history(orc)
    assert J_err < 1e-12, J_err
    assert err < 1e-9, err


def test_gradient3d_tpow_stored_history(orc):
    err, J_err = pins.gradient3d_history(orc, tpow=1.5)
    assert J_err < 1e-12, J_err
    assert err < 1e-9, err


def test_rtm3d_stored_history(orc):
    err = pins.rtm3d_history(orc)
    assert err < 1e-9, err


@pytest.mark.parametrize("nhx", [-1, 3])
@pytest.mark.parametrize("ratio", [1, 4])
def test_rtm2d_stored_history(orc, nhx, ratio):
    err = pins.rtm2d_history(orc, nhx=nhx, ratio=ratio)
    assert err < 1e-9, err


def test_rtm2d_stored_history_negative_ot(orc):
    err = pins.rtm2d_history(orc, nhx=2, ratio=2, ntneg=6)
    assert err < 1e-9, err


# measured on this problem: c11 0.0646, c13 0.316, c33 0.569
TAYLOR3D = {0: 0.0646, 1: 0.316, 2: 0.569}


@pytest.mark.parametrize("k", [0, 1, 2])
def test_gradient3d_taylor(orc, k):
    """dJ/dc along a blob in c11, c13 or c33 (central differences of the
    objective) against <g, dc>.  The reference reuses the forward operator as
    the adjoint (objFuncGradient3d_f.cu:1010,1023), which is the transpose
    only up to the coefficient matrix's position (C.D vs D.C), and injects the
    interpolated residual at every fine step while J sums coarse samples; so
    the ratio is neither 1 nor one number for all three coefficients.  What
    the mathematics fixes and this checks: the sign (a flipped residual or a
    swapped imaging term turns or scales it), the stability of the difference
    quotient in h, and the ratio, stated per coefficient to 2%."""
    r = pins.taylor3d(orc, k)
    assert abs(r[0] - r[1]) < 1e-4 * abs(r[1])
    assert abs(r[1] - TAYLOR3D[k]) < 0.02 * TAYLOR3D[k], r
