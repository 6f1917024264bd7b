"""Runs the C++ tests of the SolverBase-compatible HIP solver
(tests/cpp/test_iterative_diagonal_solver_hip.cc, built by build() /
make -C tests/cpp) on the GPU and checks that every case passed."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "cpp", "test_iterative_diagonal_solver_hip")

pytestmark = pytest.mark.gpu


def test_cpp_solver_interface():
    assert os.path.exists(BINARY), "C++ tests not built (make -C tests/cpp)"
    print("case: C++ IterativeDiagonalSolverHip tests", flush=True)
    r = subprocess.run([BINARY], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stderr
    assert "0 failures" in r.stdout
