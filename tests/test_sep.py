"""SEP header/binary IO (vtiamd.sep) against files the reference's lib/io wrote.

tests/golden/sep/ holds a dataset that the reference's randomBound program
(lib/random-boundary/randomBound.cpp with lib/io/myio.cpp, built under
oracle/_ref by tests/golden/make_golden.py) read from velin.H and wrote as
velout.H + velout.H@.  Reading it back through vtiamd.sep must give the axes
the program wrote and the same floats that vti_random_boundary makes from the
same input, bit for bit.  The writer and the boundary-file tools are checked
by round trips and against the header rules of myio.cpp:42-74.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import vtiamd.sep as sep

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_host.json")))
SEPDIR = os.path.join(HERE, "golden", "sep")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_reads_reference_written_dataset(vti):
    g = GOLD["sep"]
    h = os.path.join(HERE, "golden", g["header"])
    ax = sep.axes(h)
    assert ax == [(48, 0.0, 10.0), (40, 5.0, 12.5), (1, 0.0, 1.0)]
    assert sep.get(h, "data_format") == "native_float"
    out = sep.read(h)
    assert out.size == 48 * 40 and sha(out) == g["sha256"]
    # the input of that run, read back through its own header
    vin = sep.read(os.path.join(SEPDIR, "velin.H"))
    pad = g["pad"]
    mine = vti.random_boundary(vin.reshape(1, 40, 48), 1.0, 0.2, tuple(g["dims"]), (pad, pad, pad), g["pctG"],
                               g["seed"])
    assert np.array_equal(mine.ravel(), out)
    # partial read from an element offset
    assert np.array_equal(sep.read(h, 100, pos=37), out[37:137])


def test_header_rules(tmp_path):
    h = tmp_path / "t.H"
    h.write_text('n1=4 o1=0 d1=2\nlabel1="depth (m)" xn1=99\n#n1=7 is a comment-like token\nn1=5\n')
    # the last key= counts, also after a non-alphanumeric character ('#');
    # xn1= does not match n1=
    assert sep.get(str(h), "n1") == "5"
    assert sep.get(str(h), "label1") == "depth (m)"
    assert sep.get(str(h), "n9") is None
    h.write_text("n1=3\nfoo n1=8 bar\n#n1=7\n")
    assert sep.get(str(h), "n1") == "7"


@pytest.mark.parametrize("fmt", ["native_float", "xdr_float"])
def test_write_read_round_trip(tmp_path, fmt):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 6, 7)).astype(np.float32)
    h = str(tmp_path / "a.H")
    b = sep.write(h, a, [(7, 0.0, 1.0), (6, 10.0, 2.5), (5, -1.0, 4.0)], datapath=str(tmp_path) + "/", fmt=fmt,
                  argv=["prog", "x=1"])
    assert b == str(tmp_path / "a.H@") and os.path.getsize(b) == a.nbytes
    assert sep.axes(h) == [(7, 0.0, 1.0), (6, 10.0, 2.5), (5, -1.0, 4.0)]
    assert np.array_equal(sep.read(h).reshape(a.shape), a)
    raw = open(b, "rb").read()
    want = a.astype(">f4" if fmt == "xdr_float" else "<f4").tobytes()  # xdr: big-endian on disk
    assert raw == want
    # appending a key afterwards overrides the earlier value
    sep.to_header(h, n1=70)
    assert sep.axes(h)[0] == (70, 0.0, 1.0)


def test_moved_header_finds_binary(tmp_path):
    a = np.arange(12, dtype=np.float32)
    src = tmp_path / "src"
    src.mkdir()
    h = str(src / "m.H")
    sep.write(h, a, datapath=str(src) + "/", argv=["p"])
    dst = tmp_path / "dst"
    src.rename(dst)
    assert np.array_equal(sep.read(str(dst / "m.H")), a)


def test_boundary_file_tools_round_trip(vti, tmp_path):
    """getboundary then putboundary into a model with a wiped shell restores it."""
    nx, ny, nz, npad = 20, 18, 16, 4
    rng = np.random.default_rng(5)
    v = (2000 + 500 * rng.random((nz, ny, nx))).astype(np.float32)
    ax = [(nx, 0.0, 10.0), (ny, 0.0, 10.0), (nz, 0.0, 10.0)]
    hv = str(tmp_path / "v.H")
    sep.write(hv, v, ax, datapath=str(tmp_path) + "/", argv=["p"])
    hb = str(tmp_path / "b.H")
    b = sep.get_boundary_file(hv, npad, hb, datapath=str(tmp_path) + "/")
    assert b.size == vti.lib.vti_boundary_size(nx, ny, nz, npad)
    assert np.array_equal(sep.read(hb), b)
    wiped = v.copy()
    vti.check(vti.lib.vti_zero_boundary(vti.ptr(wiped), nx, ny, nz, npad))
    hw = str(tmp_path / "w.H")
    sep.write(hw, wiped, ax, datapath=str(tmp_path) + "/", argv=["p"])
    ho = str(tmp_path / "o.H")
    out = sep.put_boundary_file(hw, hb, npad, ho, datapath=str(tmp_path) + "/")
    assert np.array_equal(out, v.ravel())
    assert np.array_equal(sep.read(ho), v.ravel()) and sep.axes(ho) == ax
