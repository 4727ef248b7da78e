"""SEP-style files as the reference's lib/io writes and reads them.

A dataset is a text header (key=value pairs, appended over the file's life)
and a binary file named by its `in=` entry.  The rules follow
lib/io/myio.cpp and lib/io/myio.h:

- `get_s(file, key)` (myio.cpp:42-74): every `key=` that starts a line or
  follows a non-alphanumeric character counts, the last one wins; a value
  ends at the next space, or is the text between double quotes.
- `readFromHeader` (myio.h:368-392): the binary at `in=` from element `pos`
  on, native byte order for `data_format=native_float`/`native_double`,
  byte-swapped otherwise (the xdr formats).
- `write` (myio.h:269-317): the binary goes to `datapath` + the header's base
  name + "@"; the header gets the command line, then `in=` and
  `data_format=`; `to_header` (myio.h:186-206) appends `key=value` lines.

Also the boundary-file tools of lib/vti-fwi-3d/gpu (getboundary.cpp,
putboundary.cpp): the boundary shell of a 3D model packed in getBoundary order.
"""
from __future__ import annotations

import os
import sys

import numpy as np

__all__ = ["read_header", "get", "axes", "read", "write", "to_header", "get_boundary_file", "put_boundary_file"]


def _value_after(line: str, pos: int) -> tuple[str, int]:
    if pos < len(line) and line[pos] == '"':
        end = line.find('"', pos + 1)
        end = len(line) if end < 0 else end
        return line[pos + 1:end], end
    end = line.find(" ", pos)
    end = len(line) if end < 0 else end
    return line[pos:end], end


def get(path: str, key: str) -> str | None:
    """The value of `key` in the header at `path` (the last occurrence), or None."""
    k = key + "="
    val = None
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            p = line.find(k)
            while p >= 0:
                if p == 0 or not line[p - 1].isalnum():
                    val, end = _value_after(line, p + len(k))
                else:
                    end = p + len(k)
                p = line.find(k, end)
    return val


def read_header(path: str) -> dict[str, str]:
    """Every key of the header with its last value."""
    keys = []
    with open(path) as f:
        for line in f:
            for tok in line.split():
                if "=" in tok:
                    keys.append(tok.split("=", 1)[0])
    out = {}
    for k in dict.fromkeys(keys):
        if k and k.replace("_", "").isalnum():
            v = get(path, k)
            if v is not None:
                out[k] = v
    return out


def axes(path: str, ndim: int | None = None) -> list[tuple[int, float, float]]:
    """(n, o, d) of axes 1, 2, ... as the header gives them (n defaults to 1,
    o to 0, d to 1)."""
    h = read_header(path)
    nd = ndim or max([int(k[1:]) for k in h if k[:1] == "n" and k[1:].isdigit()] or [1])
    return [(int(float(h.get(f"n{i}", 1))), float(h.get(f"o{i}", 0.0)), float(h.get(f"d{i}", 1.0)))
            for i in range(1, nd + 1)]


def _binary_path(path: str) -> str:
    b = get(path, "in")
    if b is None:
        raise FileNotFoundError(f"{path}: no in= entry")
    if not os.path.exists(b):
        # a header moved with its binary: the binary next to it
        alt = os.path.join(os.path.dirname(os.path.abspath(path)), os.path.basename(b))
        if os.path.exists(alt):
            return alt
    return b


def read(path: str, n: int | None = None, pos: int = 0, dtype=np.float32) -> np.ndarray:
    """readFromHeader: n elements (all by default) from element pos on, as a
    1D array in file order (axis 1 fastest)."""
    b = _binary_path(path)
    fmt = get(path, "data_format") or "native_float"
    dt = np.dtype(dtype)
    if n is None:
        n = os.path.getsize(b) // dt.itemsize - pos
    a = np.fromfile(b, dtype=dt, count=n, offset=dt.itemsize * pos)
    if fmt not in ("native_float", "native_double"):
        a = a.byteswap()
    return a


def to_header(path: str, **kv) -> None:
    """to_header: append key=value lines."""
    with open(path, "a") as f:
        for k, v in kv.items():
            f.write(f"\n{k}={v}\n")


def write(path: str, data: np.ndarray, axes_: list[tuple[int, float, float]] | None = None,
          datapath: str | None = None, fmt: str = "native_float", argv: list[str] | None = None) -> str:
    """write + to_header: the binary to datapath + basename(path) + "@", the
    header appended with the command line, in=, data_format= and the axes.
    Returns the binary's path."""
    datapath = datapath if datapath is not None else os.path.dirname(os.path.abspath(path)) + "/"
    b = datapath + os.path.basename(path) + "@"
    a = np.ascontiguousarray(data)
    (a if fmt in ("native_float", "native_double") else a.byteswap()).tofile(b)
    with open(path, "a") as f:
        f.write(" ".join(argv if argv is not None else sys.argv) + " \n")
        f.write(f"\nin={b}\ndata_format={fmt}\n")
    for i, (n, o, d) in enumerate(axes_ or [(a.size, 0.0, 1.0)], start=1):
        to_header(path, **{f"n{i}": n, f"o{i}": o, f"d{i}": d})
    return b


def get_boundary_file(v_header: str, npad: int, out_header: str, datapath: str | None = None) -> np.ndarray:
    """getboundary.cpp: the boundary shell of the 3D model at v_header
    (n1 = nx fastest), packed in getBoundary order, written as a 1D dataset."""
    from . import _lib, check, ptr

    (nx, _, _), (ny, _, _), (nz, _, _) = axes(v_header, 3)
    v = read(v_header, nx * ny * nz)
    nb = int(_lib.vti_boundary_size(nx, ny, nz, npad))
    b = np.zeros(nb, np.float32)
    check(_lib.vti_get_boundary(ptr(b), ptr(v), nx, ny, nz, npad), "vti_get_boundary")
    write(out_header, b, [(nb, 0.0, 1.0)], datapath)
    return b


def put_boundary_file(v_header: str, boundary_header: str, npad: int, out_header: str,
                      datapath: str | None = None) -> np.ndarray:
    """putboundary.cpp: the model at v_header with its boundary shell replaced
    by the packed shell at boundary_header, written with the model's axes."""
    from . import _lib, check, ptr

    ax = axes(v_header, 3)
    (nx, _, _), (ny, _, _), (nz, _, _) = ax
    v = read(v_header, nx * ny * nz).copy()
    nb = int(_lib.vti_boundary_size(nx, ny, nz, npad))
    b = read(boundary_header, nb)
    check(_lib.vti_put_boundary(ptr(b), ptr(v), nx, ny, nz, npad), "vti_put_boundary")
    write(out_header, v, ax, datapath)
    return v
