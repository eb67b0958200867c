"""The drop-in C++ solver seams on the device, through a C++ program compiled against include/
and linked with libmetriko_hip.so (tests/native/seams_test.cpp, built by metriko_amd/build.py):
suitesparse::LinearSolver / PositiveDefiniteSolver / SquareSolver (real and complex,
Hermitian and not), solveSquare, both solveSmallestEig overloads (the smallest eigenvalue of a
grid Laplacian against its closed form), CholeskyWrapper and LMSolver with DiagonalDamping
(the Rosenbrock minimiser)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "native", "bin", "seams_test")
SEAMS = ["PositiveDefiniteSolver<double>", "PositiveDefiniteSolver<complex>", "SquareSolver<double>",
         "SquareSolver<complex>", "solveSquare", "solveSmallestEig(SprsC)", "solveSmallestEig(SprsD)",
         "CholeskyWrapper", "LMSolver+DiagonalDamping"]


@pytest.mark.gpu
def test_cpp_solver_seams():
    assert os.path.exists(EXE), "tests/native/bin/seams_test is not built (metriko_amd/build.py)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for s in SEAMS:
        assert f"ok {s} " in r.stdout, s


def test_seams_test_is_built():
    """build() compiles the C++ seams test (no GPU needed to link it)."""
    assert os.path.exists(EXE)
