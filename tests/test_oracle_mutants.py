"""The step-control pins reject every plausible misreading of the ladder.

Each ORC_MUTANT value of oracle/cochlea_oracle.cpp compiles one mistake into
the oracle (middle-band accept, the CPU ode::Tester thresholds, F tested
before L, |v| in the fault limit, a count instead of a sum, re-running a
halving step, advancing time by the old dt, a non-strict m1tol test, a summed
e_a).  The oracle is rebuilt with each one, and tests/test_oracle_ladder.py
must fail against it while passing against the unmutated build."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cochlea_oracle.cpp")
MUTANTS = range(1, 10)


def _build(k, out_dir):
    so = os.path.join(out_dir, f"liboracle_m{k}.so")
    subprocess.run(["g++", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c++17",
                    f"-DORC_MUTANT={k}", "-shared", "-o", so, SRC], check=True)
    return so


def _ladder(so):
    env = dict(os.environ, CM_ORACLE_LIB=so)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           os.path.join(ROOT, "tests", "test_oracle_ladder.py")],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)


def test_every_mutant_fails_the_ladder_pins(tmp_path):
    with ThreadPoolExecutor(4) as ex:
        libs = dict(zip([0, *MUTANTS], ex.map(lambda k: _build(k, str(tmp_path)), [0, *MUTANTS])))
    base = _ladder(libs[0])
    assert base.returncode == 0, base.stdout[-2000:]
    survivors = []
    for k in MUTANTS:
        r = _ladder(libs[k])
        if r.returncode == 0:
            survivors.append(k)
        else:
            assert "failed" in r.stdout, r.stdout[-2000:]
    assert not survivors, f"mutants not caught by the ladder pins: {survivors}"
