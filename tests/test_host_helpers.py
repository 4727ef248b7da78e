"""Host helpers of libvtiamd (CPU) against the compiled reference's outputs.

tests/golden/reference_host.json holds SHA-256 digests of what the
reference's own code produced (built from its sources under oracle/_ref by
tests/golden/make_golden.py): the random-boundary program, the parameter
maps, the 2D taper, the boundary-shell packing and the Ricker wavelet.  Every
comparison here is bit for bit.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_host.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rb_input(dims, seed):
    n1, n2, n3 = dims
    rng = np.random.default_rng(seed)
    z = np.arange(n3 * n2 * n1, dtype=np.float64).reshape(n3, n2, n1)
    return (1500.0 + 2500.0 * rng.random() + 0.01 * (z % 997)).astype(np.float32)


def test_glibc_rand_matches_libc(vti):
    libc = ctypes.CDLL("libc.so.6")
    for seed in (0, 1, 7, 2017, 123456789):
        libc.srand(seed)
        want = np.array([libc.rand() for _ in range(2000)], np.int32)
        assert np.array_equal(vti.glibc_rand(seed, 2000), want)


@pytest.mark.parametrize("case", GOLD["random_boundary"], ids=lambda c: f"{c['dims']}-{c['pctG']}-{c['seed']}")
def test_random_boundary_bitexact(vti, case):
    dims = tuple(case["dims"])
    vin = rb_input(dims, case["input_seed"])
    assert sha(vin) == case["sha256_in"]
    pad = case["pad"]
    out = vti.random_boundary(vin, 1.0, 0.2, dims, (pad, pad, pad), case["pctG"], case["seed"])
    flat = out.ravel()
    for i, v in zip(case["samples"]["index"], case["samples"]["value"]):
        assert flat[i] == np.float32(v)
    assert int(np.count_nonzero(out != vin)) == case["changed_points"]
    assert sha(out) == case["sha256_out"]


def test_random_boundary_properties(vti):
    """Interior untouched; shell factors within (eval/3, 1) as the generator bounds them."""
    dims, pad = (60, 50, 1), 10
    vin = np.full((1, 50, 60), 2000.0, np.float32)
    out = vti.random_boundary(vin, 1.0, 0.2, dims, (pad, pad, pad), 0.9, 3)
    inner = out[0, pad:-pad, pad:-pad]
    assert np.all(inner == 2000.0)
    ratio = out / vin
    assert ratio.min() > 0.2 / 3 - 1e-6 and ratio.max() <= 1.0


def test_conversions_bitexact(vti):
    g = GOLD["conversions"]
    inp = g["inputs"]
    n = inp["n"]
    rng = np.random.default_rng(inp["seed"])
    v = rng.uniform(0.4, 1.5, n).astype(np.float32)
    eps = rng.uniform(0.0, 0.3, n).astype(np.float32)
    dl = rng.uniform(-0.05, 0.2, n).astype(np.float32)
    dv = rng.standard_normal(n).astype(np.float32)
    de = rng.standard_normal(n).astype(np.float32)
    dd = rng.standard_normal(n).astype(np.float32)
    gg = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    v0, e0, d0 = inp["v0"], inp["eps0"], inp["del0"]
    L, p = vti.lib, vti.ptr
    c = [np.zeros(n, np.float32) for _ in range(6)]
    vti.check(L.vti_vepsdel_to_cij(*[p(a) for a in c[:3]], p(v), p(eps), p(dl), v0, e0, d0, n))
    assert [sha(a) for a in c[:3]] == g["VEpsDel2Cij"]
    vti.check(L.vti_dvepsdel_to_dcij(*[p(a) for a in c], p(v), p(eps), p(dl), p(dv), p(de), p(dd), v0, e0, d0, n))
    assert [sha(a) for a in c] == g["dVEpsDel2dCij"]
    out = [np.zeros(n, np.float32) for _ in range(3)]
    vti.check(L.vti_gradcij_to_gradvepsdel(*[p(a) for a in out], *[p(a) for a in gg], p(v), p(eps), p(dl), v0, e0, d0, n))
    assert [sha(a) for a in out] == g["GradCij2GradVEpsDel"]


def test_check_eps_del_bitexact(vti):
    g = GOLD["conversions"]["checkEpsDel"]
    n = g["n"]
    rng = np.random.default_rng(g["seed"])
    e = rng.uniform(0.0, 0.3, n).astype(np.float32)
    d = rng.uniform(0.0, 0.3, n).astype(np.float32)
    d[:: g["equal_every"]] = e[:: g["equal_every"]]
    e2, d2, m, count = vti.check_eps_del(e, d, g["eps0"], g["del0"])
    assert [sha(e2), sha(d2), sha(m)] == g["sha256"]
    assert count == g["projected"]
    # what the projection fixes: del0*del <= eps0*eps everywhere afterwards, equal on the projected points
    assert np.all(d2 <= e2 * (1 + 1e-6))
    assert np.allclose(d2[m == 2], e2[m == 2], rtol=1e-6)


@pytest.mark.parametrize("nx,nz,npad", [(256, 256, 40), (1024, 512, 40), (30, 20, 7)])
def test_taper2d_bitexact(vti, nx, nz, npad):
    t = vti.init_taper2d(nx, nz, npad)
    assert sha(t) == GOLD["misc"][f"init_abc2d_{nx}_{nz}_{npad}"]


def test_boundary_pack_bitexact(vti):
    m = GOLD["misc"]["getBoundary"]
    nx, ny, nz, npad = m["dims"]
    rng = np.random.default_rng(m["seed"])
    # the fixture generator drew the conversion inputs first from the same stream
    n = 4096
    for _ in range(3):
        rng.uniform(size=n)
    for _ in range(6):
        rng.standard_normal(n)
    vol = rng.standard_normal(nx * ny * nz).astype(np.float32)
    nb = vti.lib.vti_boundary_size(nx, ny, nz, npad)
    b = np.zeros(nb, np.float32)
    vti.check(vti.lib.vti_get_boundary(vti.ptr(b), vti.ptr(vol), nx, ny, nz, npad))
    assert sha(b) == m["sha256"]
    z = vol.copy()
    vti.check(vti.lib.vti_zero_boundary(vti.ptr(z), nx, ny, nz, npad))
    assert sha(z) == GOLD["misc"]["zeroBoundary"]
    # put(get(v)) into a zeroed shell restores v
    vti.check(vti.lib.vti_put_boundary(vti.ptr(b), vti.ptr(z), nx, ny, nz, npad))
    assert np.array_equal(z, vol)


@pytest.mark.parametrize(
    "args", [(15.0, 2000, 1e-3, 0.0, 0.1, 1.0), (10.0, 3001, 1e-3, 0.0, 0.12, 2.5), (12.0, 3000, 1e-3, -0.05, 0.1, 1.0)]
)
def test_ricker_bitexact(vti, args):
    freq, nt, dt, ot, tdelay, s = args
    w = vti.ricker(freq, nt, dt, ot, tdelay, s)
    assert sha(w) == GOLD["misc"][f"ricker_{freq}_{nt}_{dt}_{ot}_{tdelay}_{s}"]


def test_damping3d_closed_form(vti, orc):
    """0.95 + 0.05 cos(pi*dist/npad), dist = npad-i low side, i-n+npad high side (abc3d.cu:23-50)."""
    nx, ny, nz, npad = 37, 29, 41, 9
    d = vti.init_damping3d(nx, ny, nz, npad)
    assert np.array_equal(d, orc.init_damping3d(nx, ny, nz, npad))

    def prof(n):
        i = np.arange(n)
        dist = np.where(i < npad, npad - i, np.where(i > n - npad, i - n + npad, 0)).astype(np.float32)
        return (0.95 + (1 - 0.95) * np.cos(3.14159265359 * dist.astype(np.float64) / npad)).astype(np.float32)

    want = (prof(ny)[:, None] * prof(nx)[None, :]).astype(np.float32).ravel()
    assert np.array_equal(d[: nx * ny], want)
    assert np.array_equal(d[nx * ny:], prof(nz))
    assert d[nx * ny + nz // 2] == 1.0  # exactly one in the z interior


@pytest.mark.parametrize("nx,ny,nz", [(8192, 9000, 16), (8192, 8200, 12), (9000, 9000, 40), (572, 572, 460),
                                      (163, 140, 128), (2048, 2048, 8)])
def test_zchunk_stays_within_32bit_offsets(vti, nx, ny, nz):
    """The 3D kernels address a chunk's planes plus a 4-plane halo on each side
    with 32-bit byte offsets from the chunk's first plane.  x-y planes of
    67M-89.5M points allow only 4-plane chunks; the small-grid rule (at least
    8 planes) must not override that."""
    pitch = (nx + 63) // 64 * 64
    plane_bytes = 4 * pitch * ny
    for requested, fixed in ((64, 0), (8, 0), (64, 1), (32, 1)):
        zc = vti.lib.vti3d_zchunk_for(nx, ny, nz, requested, fixed)
        assert zc >= 4 and zc % 4 == 0
        assert (zc + 8) * plane_bytes <= 0xFFFFFFFF, (nx, ny, zc)
    if plane_bytes * 12 > 0xFFFFFFFF:
        assert vti.lib.vti3d_zchunk_for(nx, ny, nz, 64, 0) == 4
    assert vti.lib.vti3d_zchunk_for(572, 572, 460, 64, 0) == 64  # the full 460x572x572 grid keeps 64


def test_zchunk_plane_too_large(vti):
    assert vti.lib.vti3d_zchunk_for(16384, 16384, 16, 64, 0) == 0
