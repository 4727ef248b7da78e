"""The C ABI: every function include/vti_amd.h declares is exported by
libvtiamd.so, and the ctypes binding declares the same arity and argument
kinds.  Loads the library only (no GPU call)."""
import ctypes
import os
import re

import pytest

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "vti_amd.h")


def _strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def declarations():
    src = _strip_comments(open(HDR).read())
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w \*]*?)\b(vti\w*)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        ret, name, params = m.group(1).strip(), m.group(2), " ".join(m.group(3).split())
        ps = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        out[name] = (ret, ps)
    return out


def kind(decl):
    if "*" in decl:
        return "ptr"
    t = decl.rsplit(" ", 1)[0].replace("const ", "").strip()
    return {"int": "int", "float": "float", "double": "f64", "long long": "i64", "size_t": "size"}[t]


CT = {ctypes.c_int: "int", ctypes.c_float: "float", ctypes.c_double: "f64", ctypes.c_longlong: "i64",
      ctypes.c_size_t: "size",
      ctypes.c_void_p: "ptr", ctypes.c_char_p: "ptr"}


def test_header_parses():
    d = declarations()
    assert len(d) >= 30
    assert "vti3d_objfunc_gradient" in d and "vti2d_gn_hessian" in d


def test_every_declared_symbol_is_exported(vti):
    lib = ctypes.CDLL(vti.LIB_PATH)
    missing = [n for n in declarations() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", sorted(declarations()))
def test_binding_matches_header(vti, name):
    ret, params = declarations()[name]
    assert name in vti._SIGS, f"{name} has no ctypes binding"
    _, args = vti._SIGS[name]
    got = [CT.get(a, "ptr") for a in args]
    assert got == [kind(p) for p in params], (name, got, params)
