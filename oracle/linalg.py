"""Oracle: small dense / sparse linear-algebra restatements used by the IGM setup.

Test infrastructure only; see ``oracle/__init__.py``.
"""
import numpy as np


def fullpiv_lu_kernel(A):
    """Eigen 3.4 ``FullPivLU<MatrixXd>(A).kernel()`` (called at iter_rounding_init.h:129).

    Published algorithm: complete pivoting LU where step k picks the entry of
    largest magnitude in the trailing block, first in column-major order on
    ties; elimination stops when that maximum is exactly 0.  The rank counts
    pivots with |U(i,i)| > maxPivot * threshold, threshold = eps * min(rows, cols).
    Kernel: solve the leading rank x rank upper-triangular block against the
    trailing columns, negate, place rows by the column permutation Q and put
    an identity on the free columns.  A zero column is returned when the
    kernel is trivial (Eigen returns max(1, dimker) columns).
    """
    f = fullpiv_lu(A)
    return _kernel_from_lu(f)


def fullpiv_lu(A):
    """Eigen 3.4 ``FullPivLU::computeInPlace``: returns the packed LU, the row and column
    transpositions of every step (k <-> row_t[k], k <-> col_t[k]; identity after the step at
    which the trailing block is exactly zero), the number of nonzero pivots, the largest
    pivot magnitude and Q (Q[i] = the original column at position i)."""
    A = np.array(A, dtype=np.float64, copy=True)
    rows, cols = A.shape
    size = min(rows, cols)
    rowp = np.arange(rows)
    colq = np.arange(cols)
    row_t = np.arange(size)
    col_t = np.arange(size)
    nonzero_pivots = size
    max_pivot = 0.0
    lu = A
    for k in range(size):
        blk = np.abs(lu[k:, k:])
        # column-major first maximum: transpose then argmax over flattened rows
        flat = blk.T.reshape(-1)
        idx = int(np.argmax(flat))
        biggest = flat[idx]
        cj, ri = divmod(idx, rows - k)
        row_big, col_big = ri + k, cj + k
        if biggest == 0.0:
            nonzero_pivots = k
            break
        if biggest > max_pivot:
            max_pivot = biggest
        row_t[k], col_t[k] = row_big, col_big
        if k != row_big:
            lu[[k, row_big], :] = lu[[row_big, k], :]
            rowp[[k, row_big]] = rowp[[row_big, k]]
        if k != col_big:
            lu[:, [k, col_big]] = lu[:, [col_big, k]]
            colq[[k, col_big]] = colq[[col_big, k]]
        if k < rows - 1:
            lu[k + 1:, k] /= lu[k, k]
        if k < size - 1:
            lu[k + 1:, k + 1:] -= np.outer(lu[k + 1:, k], lu[k, k + 1:])
    return dict(lu=lu, row_t=row_t, col_t=col_t, nonzero_pivots=nonzero_pivots, max_pivot=max_pivot, q=colq)


def _kernel_from_lu(f):
    """Eigen 3.4 ``FullPivLU::kernel()`` from the factorisation of ``fullpiv_lu``."""
    lu, colq = f["lu"], f["q"]
    nonzero_pivots, max_pivot = f["nonzero_pivots"], f["max_pivot"]
    rows, cols = lu.shape
    size = min(rows, cols)
    threshold = np.finfo(np.float64).eps * size
    pre = max_pivot * threshold
    piv = [i for i in range(nonzero_pivots) if abs(lu[i, i]) > pre]
    rank = len(piv)
    dimker = cols - rank
    if dimker == 0:
        return np.zeros((cols, 1)), rank
    m = lu[:rank, :].copy()
    for i in range(rank):
        if i:
            m[i, :i] = 0.0
        m[i, i:] = lu[piv[i], i:]
    for i in range(rank):
        m[i, :i] = 0.0  # strictly lower part of the leading block
    for i in range(rank):
        m[:, [i, piv[i]]] = m[:, [piv[i], i]]
    # upper triangular solve in place on the trailing columns
    R = m[:, :rank]
    B = m[:, rank:]
    X = np.zeros_like(B)
    for i in range(rank - 1, -1, -1):
        X[i] = (B[i] - R[i, i + 1:] @ X[i + 1:]) / R[i, i]
    m[:, rank:] = X
    for i in range(rank - 1, -1, -1):
        m[:, [i, piv[i]]] = m[:, [piv[i], i]]
    K = np.zeros((cols, dimker))
    for i in range(rank):
        K[colq[i], :] = -m[i, rank:]
    for k in range(dimker):
        K[colq[rank + k], k] = 1.0
    return K, rank
