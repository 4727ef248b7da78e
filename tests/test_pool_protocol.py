"""The large-block reuse protocol of the library (csrc/vti_pool.h), run
against a mocked stream / event layer on the CPU (tests/pool_protocol.cpp):
no two owners of a block overlap, an event is destroyed only after it
completed and while no wait on it is pending, a block goes back to the
runtime only when idle, double releases are refused.  The previous protocol
(an event destroyed right after a stream was made to wait on it) must fail
the same checks."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "thesis-source-code_amd", "csrc")


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pool") / "pool_protocol")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-I", CSRC,
                    os.path.join(HERE, "pool_protocol.cpp"), "-o", exe], check=True, timeout=300)
    return exe


def test_pool_protocol_holds(prog):
    r = subprocess.run([prog], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "violations 0" in r.stdout


def test_previous_protocol_is_caught(prog):
    r = subprocess.run([prog, "old"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    assert "destroyed" in r.stderr
