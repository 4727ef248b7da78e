"""include/vti_compat.hpp restores the reference's C++ driver signatures.

Compiles and links (CPU only, no GPU call): the implementation unit
together with a caller that sees only the reference's lib/vti-fwi-2d/wave.h,
and checks the 3D declarations against lib/vti-fwi-3d/gpu/wave3d.h as text
(that header needs CUDA's stream type)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "thesis-source-code_amd")
REF = "/root/reference/lib"

needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")

IMPL = r"""
#include <string>
#include <type_traits>
#include <vector>
%(wave)s
%(lbfgs)s
#define VTI_COMPAT_IMPLEMENTATION
#include "vti_compat.hpp"
static_assert(std::is_same<decltype(&synthetic_f), void (*)(float *, const float *, const float *, const float *,
              const float *, const int *, int, const int *, int, const float *, int, int, int, int, float, float,
              float, float, float)>::value, "synthetic_f");
"""

CALLER = r"""
#include "%(wave)s"
#include "%(lbfgs)s"
float call(float *g, const float *d, const float *c, const float *w, const int *s, const int *r, const float *t) {
  synthetic_f(g, c, c, c, w, s, 1, r, 1, t, 10, 10, 10, 4, 1.f, 1.f, 1.f, 1.f);
  GNhessian_f(g, g, g, d, c, c, c, c, c, c, w, s, 1, r, 1, t, 10, 10, 10, 4, 1.f, 1.f, 1.f, 1.f, 0.f);
  float m[1];
  GNhessianVEpsDel(g, g, g, d, g, g, g, c, c, c, 1.f, 1.f, 1.f, w, s, 1, r, 1, t, 10, 10, 10, 4, 1.f, 1.f, 1.f, 1.f, 0.f, 0.f, m);
  objFuncGradientVEpsDel(g, g, g, d, g, g, g, 1.f, 1.f, 1.f, w, s, 1, r, 1, t, 10, 10, 10, 4, 1.f, 1.f, 1.f, 1.f, 0.f, 0.f, m);
  int m5 = 5, iflag = 0, isave[8];
  float f = 0.f, dsave[14];
  lbfgs(10, m5, g, f, g, false, g, g, iflag, isave, dsave);
  return objFuncGradient_f(g, g, g, d, c, c, c, w, s, 1, r, 1, t, 10, 10, 10, 4, 1.f, 1.f, 1.f, 1.f);
}
"""


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@needs_gxx
def test_shim_compiles_standalone(tmp_path):
    (tmp_path / "impl.cpp").write_text(IMPL % {"wave": "", "lbfgs": ""})
    _run(["g++", "-std=c++17", "-Wall", "-Werror", "-fPIC", "-c", "impl.cpp", "-I", INC, "-o", "impl.o"], tmp_path)


@needs_gxx
@needs_ref
def test_2d_callers_link_against_shim(tmp_path, vti):
    wave = os.path.join(REF, "vti-fwi-2d", "wave.h")
    lb = os.path.join(REF, "lbfgs", "lbfgs.h")
    inc = ["-I", os.path.join(REF, "lbfgs"), "-I", os.path.join(REF, "LinearSolver")]
    # lbfgs.h #defines epsilon, which breaks <cmath> after it: the caller includes it, the
    # implementation unit (as in INTEGRATION) does not, and the link checks the signature
    (tmp_path / "impl.cpp").write_text(IMPL % {"wave": f'#include "{wave}"', "lbfgs": ""})
    (tmp_path / "caller.cpp").write_text(CALLER % {"wave": wave, "lbfgs": lb})
    _run(["g++", "-std=c++17", "-fPIC", "-c", "impl.cpp", "-I", INC, *inc, "-o", "impl.o"], tmp_path)
    _run(["g++", "-std=c++17", "-fPIC", "-c", "caller.cpp", *inc, "-o", "caller.o"], tmp_path)
    _run(["g++", "-shared", "impl.o", "caller.o", "-L", LIBDIR, "-lvtiamd", "-Wl,--no-undefined",
          "-Wl,-rpath-link,/opt/rocm/lib", "-o", "libcaller.so"], tmp_path)


def _decl_types(text, name):
    m = re.search(r"\b(void|float)\s+" + name + r"\s*\(([^)]*)\)\s*;", text, flags=re.S)
    assert m, name
    types = []
    for p in m.group(2).split(","):
        p = " ".join(p.replace("&", " & ").replace("*", " * ").split())
        p = re.sub(r"\bstd::", "", p)
        toks = p.split(" ")
        if toks[-1] not in ("*", "&") and len(toks) > 1:
            toks = toks[:-1]  # drop the parameter name
        types.append(" ".join(toks))
    return m.group(1), types


@needs_ref
@pytest.mark.parametrize("name", ["objFuncGradientCij3d_f", "objFuncCij3d_f", "modelDataCij3d_f", "rtmCij3d_f"])
def test_3d_declarations_match_reference(name):
    ref = open(os.path.join(REF, "vti-fwi-3d", "gpu", "wave3d.h")).read()
    ours = open(os.path.join(INC, "vti_compat.hpp")).read()
    assert _decl_types(ours, name) == _decl_types(ref, name)
