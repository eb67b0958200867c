"""Pins of the constraint kernel basis: Eigen 3.4 FullPivLU<MatrixXd>(CSmall).kernel()
(iter_rounding_init.h:129-130), in the oracle (oracle/linalg.py) and in the drop-in header
(include/metriko/core/solver/matrix_ops.h, compiled here with g++).

* Pivot sequence and factors against LAPACK's complete-pivoting LU (dgetc2) on random real
  matrices, where no two candidates tie (dgetc2 takes the last maximum on ties, Eigen the
  first in column-major order, so the tie rule is pinned by the hand-worked cases instead).
* Hand-worked kernels on small integer matrices whose pivots are forced by ties.
* The basis is the reduced one: identity on the free columns Q[rank..] in that order, minus
  U11^-1 U12 on the pivot columns; checked against a direct solve for random integer matrices
  with ties, and header == oracle on them (pivots and layout exactly, values to round-off: the
  back substitution's summation order, like that of Eigen's triangular solve, is not pinned).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.linalg import lapack

from oracle.linalg import fullpiv_lu, fullpiv_lu_kernel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("seed", range(10))
def test_pivots_match_lapack_complete_pivoting(seed):
    g = np.random.default_rng(seed)
    n = int(g.integers(3, 12))
    A = g.standard_normal((n, n))
    lu, ipiv, jpiv, info = lapack.dgetc2(A)
    assert info == 0
    f = fullpiv_lu(A)
    assert np.array_equal(f["row_t"], ipiv) and np.array_equal(f["col_t"], jpiv)
    assert np.abs(f["lu"] - lu).max() <= 1e-12 * np.abs(lu).max()


KATS = [
    # (A, rank, K): derivations in the docstring of test_hand_worked_kernels
    (np.array([[1., 2., 3.], [2., 4., 6.]]), 1,
     np.array([[0., 1.], [1., 0.], [-2. / 3., -1. / 3.]])),
    (np.array([[1., -1., 0., 0.], [0., 1., -1., 0.], [-1., 0., 1., 0.]]), 2,
     np.array([[1., 0.], [1., 0.], [1., 0.], [0., 1.]])),
    (np.array([[0., 2.], [2., 1.]]), 2, np.zeros((2, 1))),
]


def test_hand_worked_kernels():
    """(1) [[1,2,3],[2,4,6]]: the column-major scan with strict > ends at 6 = (1,2); rows 0/1
    and columns 0/2 swap, Q = [2,1,0]; the trailing block [0,0] stops the elimination; rank 1;
    U = [6,4,2] gives -U11^-1 U12 = [-2/3, -1/3] on Q[0] = column 2, and the free columns
    Q[1] = 1, Q[2] = 0 take the identity in that order.
    (2) The 3-cycle of differences [[1,-1,0,0],[0,1,-1,0],[-1,0,1,0]]: all entries tie at 1,
    so (0,0) then (1,1) pivot with no swaps; the last row cancels; rank 2; back substitution
    gives -[-1,-1] on columns 0, 1 for free column 2, and column 3 (empty) is its own vector.
    (3) [[0,2],[2,1]]: (1,0) beats the 0 at (0,0) and ties the 2 at (0,1), which comes later
    in column-major order: a row swap, full rank, a single zero column is returned."""
    for A, rank, K in KATS:
        Kc, r = fullpiv_lu_kernel(A)
        assert r == rank
        np.testing.assert_allclose(Kc, K, atol=1e-15, rtol=0)


def _reduced_basis_check(A, K, rank, q):
    """K restricted to the free columns q[rank:] is the identity (in that order) and A K = 0."""
    cols = A.shape[1]
    if cols == rank:
        assert np.abs(K).max() == 0.0
        return
    np.testing.assert_array_equal(K[q[rank:], :], np.eye(cols - rank))
    assert np.abs(A @ K).max() <= 1e-12 * max(1.0, np.abs(A).max())
    # the pivot-column rows are the unique solution of A[:, piv] X = -A[:, free]
    piv = q[:rank]
    X = np.linalg.lstsq(A[:, piv], -A[:, q[rank:]], rcond=None)[0]
    np.testing.assert_allclose(K[piv, :], X, atol=1e-10)


@pytest.mark.parametrize("seed", range(8))
def test_kernel_is_the_reduced_basis(seed):
    g = np.random.default_rng(seed)
    r, c, rank = int(g.integers(2, 9)), int(g.integers(4, 14)), int(g.integers(1, 4))
    A = (g.integers(-1, 2, (r, rank)) @ g.integers(-1, 2, (rank, c))).astype(float)
    K, rk = fullpiv_lu_kernel(A)
    assert rk == np.linalg.matrix_rank(A)
    _reduced_basis_check(A, K, rk, fullpiv_lu(A)["q"])


@pytest.fixture(scope="module")
def kernel_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("native") / "fullpiv_kernel"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "fullpiv_kernel.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def _native_kernels(exe, mats):
    lines = []
    for A in mats:
        lines.append(f"{A.shape[0]} {A.shape[1]}")
        lines.append(" ".join(repr(float(x)) for x in A.reshape(-1)))
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True).stdout
    it = iter(out.strip().splitlines())
    res = []
    for ln in it:
        _, rank, kr, kc = ln.split()
        K = np.array([[float(x) for x in next(it).split()] for _ in range(int(kr))]).reshape(int(kr), int(kc))
        res.append((int(rank), K))
    return res


def test_header_matches_oracle(kernel_exe):
    g = np.random.default_rng(11)
    mats = [A for A, _, _ in KATS]
    for _ in range(30):   # CSmall-like: small integers, many ties
        r, c, rank = int(g.integers(2, 12)), int(g.integers(3, 20)), int(g.integers(1, 6))
        mats.append((g.integers(-1, 2, (r, rank)) @ g.integers(-1, 2, (rank, c))).astype(float))
    for A, (rank, K) in zip(mats, _native_kernels(kernel_exe, mats)):
        Ko, ro = fullpiv_lu_kernel(A)
        assert rank == ro
        # same pivots and layout: the identity rows of the free columns are exact; the back
        # substitution sums in another order, so the pivot-column rows agree to round-off
        free = fullpiv_lu(A)["q"][ro:]
        assert np.array_equal(K[free], Ko[free])
        np.testing.assert_allclose(K, Ko, atol=1e-14, rtol=0)
