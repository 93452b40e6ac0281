"""The C++ policy set of include/micm_hip_policies.hpp (HipProcessSet, HipState,
HipRosenbrockSolver, HipSolverBuilder) plugged into the reference's own SolverBuilder: the
README example and 1000- and 300-cell versions of it, compiled against the reference's
headers (oracle/_ref/hip_builder_test, tests/cpp/hip_solver_builder_test.cpp), give the
reference CPU solver's results bit for bit on the same vector-ordered policies."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "hip_builder_test")

# README.md:171-181, the reference's printed table
README_TABLE = [
    "0,   1.18e+01,   5.90e+00,   1.91e+00",
    "500,   6.79e+00,   9.05e+00,   3.32e+00",
    "1000,   3.83e+00,   1.07e+01,   4.21e+00",
    "4500,   6.13e-02,   1.28e+01,   5.41e+00",
]


def test_cpp_policy_set_matches_the_reference_solver():
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/hip_builder_test is built only where the reference headers are present")
    print("C++ policy set: running", EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.count(": ok") == 4 and "MISMATCH" not in r.stdout
    rows = [" ".join(line.split()) for line in r.stdout.splitlines()]
    for row in README_TABLE:
        assert " ".join(row.split()) in rows
