"""The request server's admission rule (csrc/vti_serve_policy.h) in a
discrete-event simulation of its workers on the CPU (tests/serve_policy.cpp):
requests start in arrival order, a 3D modeling request never runs beside
another run when 3D requests run alone, the estimates in flight stay within
the budget, and every request completes.  With 3D requests sharing the
device the simulation must show overlapping 3D runs, so that the overlap
check is seen to bite."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "thesis-source-code_amd", "csrc")


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("serve") / "serve_policy")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-I", CSRC,
                    os.path.join(HERE, "serve_policy.cpp"), "-o", exe], check=True, timeout=300)
    return exe


def _overlaps(out):
    return int(re.search(r"overlapping 3D runs (\d+)", out).group(1))


def test_exclusive_3d_rule_holds(prog):
    r = subprocess.run([prog], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "violations 0" in r.stdout
    assert _overlaps(r.stdout) == 0


def test_shared_3d_overlaps_are_seen(prog):
    r = subprocess.run([prog, "shared"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert _overlaps(r.stdout) > 0
