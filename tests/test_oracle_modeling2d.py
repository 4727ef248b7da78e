"""The oracle's modeling_f (lib/vti-fwi-2d/modeling_f.cu:10-86) pinned against
its synthetic_f, which the plane-wave and transpose tests pin, through two
identities of the equations:

- isotropic (c11 = c13 = c33): p and q receive the same update and the same
  source, so p = q and synthetic_f's receiver value 1/2 (p + q) is p.  With a
  receiver on every interior point and the wavelet w'[0] = w[0],
  w'[k] = w[k] / dt^2 (synthetic_f injects dt^2 w'[it-1], modeling_f w[it-1]
  and dt^2 w[0] at level 1) every recorded slot of modeling_f is
  synthetic_f's row;
- x-only (c13 = c33 = 0): q never leaves the source point, so away from it
  1/2 (p + q) = p / 2 and modeling_f's p is twice synthetic_f's row.

The wavelet scaling is rounded in float by synthetic_f, so the identities
hold to about 1e-7; the tolerance is 1e-6.  Injecting dt^2 w, w[it] for
w[it-1], recording 1/2 (p + q) or recording before abc fails one of them."""
import numpy as np
import pytest

from common3d import rel, ricker

NX = NZ = 24
NPAD = 6
NNX, NNZ = NX + 2 * NPAD, NZ + 2 * NPAD
DX, DT = 10.0, 1e-3


def _setup(vti, nt):
    taper = vti.init_taper2d(NX, NZ, NPAD)
    w = ricker(nt, DT, 25.0, 0.03)
    ix, iz = NNX // 2, NNZ // 2
    rec = [(x, z) for z in range(4, NNZ - 4) for x in range(4, NNX - 4)]
    return taper, w, ix, iz, np.array(rec, np.int32)


def _interior(a):
    return a[..., 4:NNZ - 4, 4:NNX - 4]


def _scaled(w):
    ws = (w.astype(np.float64) / (np.float32(DT) * np.float32(DT))).astype(np.float32)
    ws[0] = w[0]
    return ws


def _pair(orc, vti, c11, c13, c33, rate, nt):
    taper, w, ix, iz, rec = _setup(vti, nt)
    m = orc.modeling2d(orc.F64, c11, c13, c33, w, ix + iz * NNX, taper, NX, NZ, nt, NPAD, DX, DX, DT, rate)
    s = orc.synthetic2d(orc.F64, c11, c13, c33, _scaled(w), np.array([[ix, iz, len(rec), 0]], np.int32), rec, taper,
                        NX, NZ, nt, NPAD, DX, DX, DT, rate)
    return m, s.reshape(s.shape[0], NNZ - 8, NNX - 8), (ix, iz)


def _v():
    v = np.full((NNZ, NNX), 1500.0, np.float32)
    v[NNZ // 2 + 5:, :] = 2200.0
    return v


@pytest.mark.parametrize("rate", [1e-3, 2e-3])
def test_isotropic_modeling_equals_synthetic_rows(orc, vti, rate):
    c = (_v() * _v()).astype(np.float32)
    nt = 121
    m, s, _ = _pair(orc, vti, c, c, c, rate, nt)
    assert m.shape[0] == s.shape[0]
    assert np.abs(m[0]).max() == 0  # slot 0 is never recorded
    for k in range(1, m.shape[0]):
        assert rel(_interior(m[k]), s[k]) < 1e-6, k
    # the wavefield reached the tapered rim, where abc acts before recording
    rim = np.abs(m[-1][:NPAD, :]).max()
    assert rim > 1e-3 * np.abs(m[-1]).max()
    # outside the updated interior p stays 0
    assert np.abs(m[:, :4, :]).max() == 0 and np.abs(m[:, :, :4]).max() == 0


def test_x_only_modeling_is_p_not_half_sum(orc, vti):
    c11 = (_v() * _v()).astype(np.float32)
    zero = np.zeros_like(c11)
    nt = 121
    m, s, (ix, iz) = _pair(orc, vti, c11, zero, zero, 1e-3, nt)
    mask = np.ones((NNZ - 8, NNX - 8), bool)
    mask[iz - 4, ix - 4] = False
    for k in range(1, m.shape[0]):
        mk = _interior(m[k])
        assert rel(mk[mask], 2 * s[k][mask]) < 1e-6, k
    # the field stays on the source row: x-only propagation
    assert np.abs(m[-1][iz + 1:, :]).max() == 0 and np.abs(m[-1][:iz, :]).max() == 0


def test_unrecorded_slots_keep_their_values(orc, vti):
    c = (_v() * _v()).astype(np.float32)
    taper, w, ix, iz, _ = _setup(vti, 41)
    init = np.full((11, NNZ, NNX), 7.0)
    m = orc.modeling2d(orc.F64, c, c, c, w, ix + iz * NNX, taper, NX, NZ, 41, NPAD, DX, DX, DT, 4e-3, init=init)
    assert np.all(m[0] == 7.0)  # level 0 is never recorded
    for k in range(1, 11):
        assert not np.any(m[k] == 7.0), k  # recorded: p everywhere, 0 outside the interior
