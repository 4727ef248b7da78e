"""include/vti_compat.hpp restores the reference's C++ driver signatures.

Compiles and links (CPU only, no GPU call): the implementation unit
together with a caller that sees only the reference's lib/vti-fwi-2d/wave.h,
and with a 3D caller that sees only the prototypes lib/vti-fwi-3d/gpu/wave3d.h
declares for the functions it calls (the header as a whole needs CUDA's
stream type), and checks the 3D declarations against that header."""
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


def _decls(text, name):
    """Every declaration of `name` (return type, parameter types)."""
    out = []
    for m in re.finditer(r"\b(void|float|double)\s+" + name + r"\s*\(([^)]*)\)\s*;", text, flags=re.S):
        types = []
        for p in m.group(2).split(","):
            p = " ".join(p.split("=")[0].replace("&", " & ").replace("*", " * ").split())
            p = re.sub(r"\bstd::", "", p)
            toks = p.split(" ")
            if toks[-1] not in ("*", "&") and len(toks) > 1:
                toks = toks[:-1]  # drop the parameter name
            types.append(" ".join(toks))
        out.append((m.group(1), types))
    return out


NAMES3D = ["objFuncGradientCij3d_f", "objFuncCij3d_f", "modelDataCij3d_f", "rtmCij3d_f", "rtmVEpsDel",
           "objFuncGradientCij3d", "objFuncCij3d", "objFuncGradientVEpsDel", "modelData3d_f"]


@needs_ref
@pytest.mark.parametrize("name", NAMES3D)
def test_3d_declarations_match_reference(name):
    """Each 3D function the shim defines has a declaration in wave3d.h with the
    same return and parameter types (objFuncGradientVEpsDel and modelData3d_f
    also have other overloads, in wave.h and wave3d.h:84)."""
    ref = _decls(open(os.path.join(REF, "vti-fwi-3d", "gpu", "wave3d.h")).read(), name)
    ours = _decls(open(os.path.join(INC, "vti_compat.hpp")).read(), name)
    ours3d = [d for d in ours if name != "objFuncGradientVEpsDel" or "vector<int> &" in d[1]]
    assert ours3d, name
    for d in ours3d:
        assert d in ref, (name, d)


CALLER3D = r"""
#include <string>
#include <vector>
%(protos)s
// the calls of fwi3d-lbfgsc.cu:113, rtm3d.cu:75, objFuncGradient3d-cij.cu:78,89 and ModelingOp3d.cu:13
double call3d(float *g, float *s, float *r, float *w, float *m, std::vector<int> &shots, std::vector<int> &gpus) {
  double f = objFuncGradientVEpsDel(g, g, g, s, 1, shots, r, w, m, m, m, m, m, 20, 20, 20, 66, 5,
                                    0.f, 0.f, 0.f, 0.f, 16.f, 16.f, 16.f, 1e-3f, 4e-3f, 1.f, 1.f, 0.f, gpus, 2);
  rtmVEpsDel(g, s, 1, shots, r, w, m, m, m, m, 20, 20, 20, 66, 5, 0.f, 0.f, 0.f, 0.f, 16.f, 16.f, 16.f, 1e-3f, 4e-3f, 0.f);
  float of = 0.f;
  objFuncCij3d(&of, s, 1, shots, r, w, m, m, 20, 20, 20, 66, 5, 0.f, 0.f, 0.f, 0.f, 16.f, 16.f, 16.f, 1e-3f, 4e-3f, gpus, 2);
  objFuncGradientCij3d(g, s, 1, shots, r, w, m, m, m, 20, 20, 20, 66, 5, 0.f, 0.f, 0.f, 0.f, 16.f, 16.f, 16.f, 1e-3f,
                       4e-3f, gpus, 2);
  modelData3d_f(g, s, 1, shots, r, w, m, m, m, 20, 20, 20, 66, 5, 0.f, 0.f, 0.f, 0.f, 16.f, 16.f, 16.f, 1e-3f, 4e-3f,
                gpus, 2);
  objFuncGradientCij3d_f(g, s, 1, shots, r, w, m, m, m, 20, 20, 20, 66, 5, 0.f, 0.f, 0.f, 0.f, 16.f, 16.f, 16.f, 1e-3f,
                         4e-3f, gpus);
  return f + of;
}
"""


@needs_gxx
@needs_ref
def test_3d_callers_link_against_shim(tmp_path, vti):
    """A caller that sees only the reference's own 3D prototypes (the lines of
    wave3d.h that declare the functions it calls) links against the shim and
    libvtiamd with --no-undefined."""
    ref = open(os.path.join(REF, "vti-fwi-3d", "gpu", "wave3d.h")).read()
    protos = []
    for name in NAMES3D:
        for m in re.finditer(r"^(?:void|float|double)\s+" + name + r"\s*\([^)]*\)\s*;", ref, flags=re.M | re.S):
            protos.append(m.group(0).replace("std::vector", "std::vector"))
    (tmp_path / "impl.cpp").write_text(IMPL % {"wave": "", "lbfgs": ""})
    (tmp_path / "caller3d.cpp").write_text(CALLER3D % {"protos": "\n".join(protos)})
    _run(["g++", "-std=c++17", "-fPIC", "-c", "impl.cpp", "-I", INC, "-o", "impl.o"], tmp_path)
    _run(["g++", "-std=c++17", "-fPIC", "-c", "caller3d.cpp", "-o", "caller3d.o"], tmp_path)
    _run(["g++", "-shared", "impl.o", "caller3d.o", "-L", LIBDIR, "-lvtiamd", "-Wl,--no-undefined",
          "-Wl,-rpath-link,/opt/rocm/lib", "-o", "libcaller3d.so"], tmp_path)
