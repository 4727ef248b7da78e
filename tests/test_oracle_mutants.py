"""Mutation test of the oracle's pins (CPU).

Each mutant is a build of the oracle with one plausible mistake planted
(-DORC_MUTANT=k, the hooks in oracle/orc_common.h): swapped 2/3-1/3 receiver
weights, swapped gc13 terms, a flipped residual sign, a wrong image weight, a
wrong t^p exponent, a shifted time index, a reversed slice interpolation.
Every pin of tests/pins.py must hold for the real oracle, and every mutant must
fail at least one of them -- otherwise that part of the oracle would be pinned
by nothing but itself.
"""
import concurrent.futures as cf
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

MUTANTS = {
    1: "3D recordData weights swapped (1/3 sx + 2/3 sz)",
    2: "3D injectResidual weights swapped",
    3: "3D first adjoint level weights swapped",
    4: "gc13 terms swapped (lx*ty + lz*tz)",
    5: "residual sign flipped (d_obs - d)",
    6: "3D RTM image receiver weights swapped",
    7: "t^p gain with exponent p+1",
    8: "residual interpolated at fine index it instead of it+1",
    9: "2D image drops the q term of 0.5 (p + q)",
    10: "2D slice interpolation weights reversed",
    11: "ODCIG offset applied to the wrong side",
    12: "gc33 imaged against ty instead of tz",
}


def _errors(lib):
    env = dict(os.environ)
    if lib:
        env["VTI_ORACLE_LIB"] = lib
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([sys.executable, os.path.join(HERE, "pins.py")], env=env, cwd=HERE, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.fixture(scope="module")
def mutant_errors(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")

    def build_and_run(k):
        out = str(d / f"liboracle_m{k}.so")
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mutant", f"K={k}", f"OUT={out}"],
                       check=True, capture_output=True, timeout=300)
        return k, _errors(out)

    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build_and_run, MUTANTS))


def test_real_oracle_meets_every_pin():
    for name, (err, bar) in _errors(None).items():
        assert err < bar, (name, err)


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_mutant_is_caught(mutant_errors, k):
    failed = [name for name, (err, bar) in mutant_errors[k].items() if not err < bar]
    assert failed, f"mutant {k} ({MUTANTS[k]}) passes every pin"
