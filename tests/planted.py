"""Copies of the oracle with one planted mistake (CGS_ORACLE_PLANT, see
oracle/cgs_oracle.hpp), for showing that a pin of the oracle fails on a
plausible error.  A check runs in a child process that loads the copy through
CGS_ORACLE_LIB; the copy lives in a temporary directory and is never used by
anything else."""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import conftest

_DIR = tempfile.mkdtemp(prefix="cgs_planted_")
_BUILT: dict[int, str] = {}


def build(plant: int) -> str:
    """The oracle built with CGS_ORACLE_PLANT = plant (-O1: the planted copy
    only has to run the check, and IEEE results do not depend on -O)."""
    if plant not in _BUILT:
        out = os.path.join(_DIR, f"liboracle_plant{plant}.so")
        src = os.path.join(conftest.ROOT, "oracle", "cgs_oracle.cpp")
        subprocess.run(["g++", "-O1", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-std=c++20", "-fPIC",
                        f"-DCGS_ORACLE_PLANT={plant}", "-shared", "-o", out, src, "-lpthread"], check=True)
        _BUILT[plant] = out
    return _BUILT[plant]


def check_fails(plant: int, module: str, func: str, *args) -> str:
    """Runs tests/<module>.py's `func(oracle, *args)` against the planted copy
    in a child process; returns its output and asserts that the check failed
    with an AssertionError (not a crash)."""
    lib = build(plant)
    code = (f"import sys; sys.path[:0] = [{os.path.join(conftest.ROOT, 'tests')!r}, {os.path.join(conftest.ROOT, 'oracle')!r}]\n"
            f"import cgs_oracle, {module}\n"
            f"try:\n    {module}.{func}(cgs_oracle, *{args!r})\n"
            f"except AssertionError as e:\n    print('PLANT CAUGHT:', str(e)[:200]); sys.exit(3)\n"
            f"print('PLANT MISSED'); sys.exit(0)\n")
    env = dict(os.environ, CGS_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300)
    assert r.returncode == 3, f"planted mistake {plant} not caught by {module}.{func}:\n{r.stdout[-2000:]}"
    return r.stdout
