"""Pins of the independent-row choice of reduce_to_linearly_independent (solver/matrix_ops.h:12-28):
COLAMD's column order of C^T and SparseQR's rank decisions in that order.

* The oracle's COLAMD (``oracle.colamd.colamd``) against SuperLU's copy of COLAMD: scipy's
  ``splu`` column permutation is COLAMD's order followed by a postorder of the column
  elimination tree of the reordered matrix (SuperLU's sp_preorder), so the oracle's order,
  put through the same postorder, must give the same permutation.  SuperLU ships a later
  COLAMD release whose dense-row / dense-column thresholds are max(16, 10 sqrt(n)) instead of
  Eigen 3.4's 0.5 n; matrices with no dense row or column under either rule pin the shared
  algorithm, and matrices with dense rows pin the dense handling under SuperLU's rule.
* Hand-worked orderings on two small patterns (derivations in the docstrings).
* SparseQR's decisions against brute force: a column is kept exactly when it raises the rank
  of the kept set (numpy matrix_rank), in the visiting order.
* The drop-in C++ header (include/metriko/core/solver/colamd.h, matrix_ops.h), compiled here
  with g++, against the oracle on random constraint-like matrices and on the Cfull matrices
  of the oracle pipeline.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.colamd import colamd, reduce_to_linearly_independent, sparse_qr_kept_columns, structural_product

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- SuperLU's column preorder (sp_coletree + TreePostorder), to compare with splu.perm_c ----
def _column_etree(A):
    """Column elimination tree (the elimination tree of A^T A) of a CSC matrix; roots point to n."""
    m, n = A.shape
    firstcol = np.full(m, n)
    for j in range(n):
        rows = A.indices[A.indptr[j]:A.indptr[j + 1]]
        firstcol[rows] = np.minimum(firstcol[rows], j)
    parent = [n] * n
    uf = list(range(n))
    root = list(range(n))

    def find(x):
        while uf[x] != x:
            uf[x] = uf[uf[x]]
            x = uf[x]
        return x

    for col in range(n):
        cset = col
        root[cset] = col
        for k in range(A.indptr[col], A.indptr[col + 1]):
            rr = int(firstcol[A.indices[k]])
            if rr >= col:
                continue
            rset = find(rr)
            rroot = root[rset]
            if rroot != col:
                parent[rroot] = col
                uf[rset] = cset
                cset = find(cset)
                root[cset] = col
    return parent


def _postorder(parent, n):
    """Depth-first postorder from the virtual root n, children in increasing order."""
    kids = [[] for _ in range(n + 1)]
    for v in range(n):
        kids[parent[v]].append(v)
    post = [0] * (n + 1)
    num = 0
    stack = [(n, 0)]
    while stack:
        v, i = stack.pop()
        if i < len(kids[v]):
            stack.append((v, i + 1))
            stack.append((kids[v][i], 0))
        else:
            post[v] = num
            num += 1
    return post


def _superlu_perm_c(A, **kw):
    A = sp.csc_matrix(A)
    A.sort_indices()
    m, n = A.shape
    p = colamd(m, n, A.indptr, A.indices, **kw)
    pos = np.empty(n, dtype=np.int64)
    pos[p] = np.arange(n)
    AC = sp.csc_matrix(A[:, p])
    AC.sort_indices()
    post = _postorder(_column_etree(AC), n)
    return np.array([post[pos[i]] for i in range(n)])


def _random_square(rng, n, density, dense_row=False):
    R = sp.random(n, n, density=density, random_state=int(rng.integers(1 << 30)), format="lil")
    if dense_row:
        r = int(rng.integers(n))
        for j in rng.choice(n, size=int(0.7 * n), replace=False):
            R[r, j] = 1.0
    A = (R.tocsc() + 4.0 * sp.identity(n)).tocsc()
    A.sort_indices()
    return A


def test_colamd_matches_superlu():
    rng = np.random.default_rng(0)
    tested = 0
    for _ in range(240):
        n = int(rng.integers(5, 120))
        A = _random_square(rng, n, rng.uniform(0.01, 0.08))
        lim = min(0.5 * n, max(16, 10 * np.sqrt(n)))
        if np.diff(A.tocsr().indptr).max() > lim or np.diff(A.indptr).max() > lim:
            continue  # a dense row or column under one of the two rules
        try:
            lu = spla.splu(A, permc_spec="COLAMD")
        except RuntimeError:
            continue
        assert np.array_equal(_superlu_perm_c(A), lu.perm_c), (n, A.nnz)
        tested += 1
    assert tested > 150


def test_colamd_dense_rows_match_superlu_rule():
    """The dense-row handling, with SuperLU's thresholds (knobs 10, sqrt rule)."""
    rng = np.random.default_rng(1)
    tested = 0
    for _ in range(80):
        n = int(rng.integers(10, 300))
        A = _random_square(rng, n, 0.05, dense_row=True)
        try:
            lu = spla.splu(A, permc_spec="COLAMD")
        except RuntimeError:
            continue
        got = _superlu_perm_c(A, dense_row=10, dense_col=10, dense_rule="sqrt")
        assert np.array_equal(got, lu.perm_c), n
        tested += 1
    assert tested > 60


def _pattern(cols, n_row):
    ptr, ind = [0], []
    for c in cols:
        ind += c
        ptr.append(len(ind))
    return n_row, len(cols), ptr, ind


def test_colamd_known_answer_dense_column():
    """4 x 4, columns {0,1,2,3}, {0}, {1}, {2,3}.  Column 0 has 4 > 0.5*4 entries: dense,
    ordered last.  The others score 0 (every row then has one live column); the degree list
    of score 0 holds 1, 2, 3 (low indices first), each pivot row is empty: p = [1, 2, 3, 0]."""
    assert colamd(*_pattern([[0, 1, 2, 3], [0], [1], [2, 3]], 4)) == [1, 2, 3, 0]


def test_colamd_known_answer_supercolumn_and_mass_elimination():
    """6 x 6, columns c0..c5 = {0}, {0,1}, {0,1}, {1,2}, {2,3}, {4}; row 5 is empty (killed).
    Row degrees 3, 3, 2, 1, 1; scores (sum of degree - 1): c0 2, c1 4, c2 4, c3 3, c4 1, c5 0.
    Step 1: c5 (score 0), empty pivot row.  Step 2: c4 (score 1), pivot row {c3}; row 1's
    set difference is 2, so c3 scores 2 + 1 - 1 = 2 and goes in front of c0.  Step 3: c3,
    pivot row {c1, c2}; both keep only row 0 (set difference 1): identical, c1 is absorbed
    into c2 (thickness 2), score min(1 + 2 - 2, 6 - 3 - 2) = 1.  Step 4: c2, pivot row {c0},
    whose only row dies: c0 is eliminated at once (order 5).  c1 takes c2's place (3) and c2
    moves to 4: orders c0..c5 = 5, 3, 4, 2, 1, 0."""
    assert colamd(*_pattern([[0], [0, 1], [0, 1], [1, 2], [2, 3], [4]], 6)) == [5, 4, 3, 1, 2, 0]


@pytest.mark.parametrize("seed", range(8))
def test_sparse_qr_keeps_rank_raising_columns(seed):
    """SparseQR's rejections against brute force: in the visiting order, a column is kept
    exactly when it is independent of the kept ones (the rank grows)."""
    g = np.random.default_rng(seed)
    m, r, n = 14, 5, 11
    M = (g.integers(-2, 3, (m, r)) @ g.integers(-1, 2, (r, n))).astype(float)
    M[:, 3] = 0.0                  # an empty column
    M[:, 7] = M[:, 1] - 2 * M[:, 4]  # a planted dependency
    order = list(g.permutation(n))
    kept = sparse_qr_kept_columns(M, order)
    expect, rank = [], 0
    for j in order:
        if np.linalg.matrix_rank(M[:, expect + [j]]) > rank:
            expect.append(j)
            rank += 1
    assert kept == expect
    assert len(kept) == np.linalg.matrix_rank(M)


def test_sparse_qr_threshold():
    """|beta| below 20 (m + n) max||col|| eps is a rejection, above it a column is kept."""
    m, n = 6, 3
    eps = np.finfo(float).eps
    a = np.array([1.0, 2.0, 0.0, -1.0, 0.0, 1.0])
    b = np.array([0.0, 1.0, 1.0, 0.0, 2.0, 0.0])
    thr = 20 * (m + n) * np.linalg.norm(b) * eps
    e = np.zeros(m)
    e[2] = 1.0
    e -= (e @ a) / (a @ a) * a + 0.0
    e -= (e @ b) / (b @ b) * b
    e /= np.linalg.norm(e)
    small = a + b + 0.1 * thr * e        # distance 0.1 thr from span(a, b): rejected
    large = a + b + 10.0 * thr * e       # distance 10 thr: kept
    assert sparse_qr_kept_columns(np.stack([a, b, small], 1), [0, 1, 2]) == [0, 1]
    assert sparse_qr_kept_columns(np.stack([a, b, large], 1), [0, 1, 2]) == [0, 1, 2]


def test_structural_product_keeps_cancellations():
    """Eigen's default sparse product stores every reached entry: x - x is an explicit zero."""
    A = sp.csr_matrix(np.array([[1.0, 1.0, 0.0], [0.0, 2.0, 1.0]]))
    B = sp.csr_matrix(np.array([[1.0, 0.0], [-1.0, 3.0], [0.0, 1.0]]))
    P = structural_product(A, B)
    assert P.nnz == 4 and P[0, 0] == 0.0 and P[0, 1] == 3.0
    assert list(P.indices[:2]) == [0, 1]


# ---- the drop-in C++ header against the oracle ---------------------------------------------
@pytest.fixture(scope="module")
def reduce_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("native") / "reduce_rows"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "reduce_rows.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def _run_native(exe, mats):
    lines = []
    for C in mats:
        C = sp.csr_matrix(C)
        lines.append(f"{C.shape[0]} {C.shape[1]} {C.nnz}")
        lines.append(" ".join(map(str, C.indptr)))
        lines.append(" ".join(map(str, C.indices)))
        lines.append(" ".join(repr(float(x)) for x in C.data))
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True).stdout
    res = []
    it = iter(out.strip().splitlines())
    for ln in it:
        order = [int(x) for x in ln.split()[1:]]
        kept = [int(x) for x in next(it).split()[1:]]
        res.append((order, kept))
    return res


def _constraint_like(rng, nrows, ncols):
    """Rows of +-1 / +-2 entries on a few columns, with planted dependent rows, negated copies
    and explicit zeros (the shapes Cfull = constraint * uncompress takes)."""
    rows = []
    for i in range(nrows):
        k = int(rng.integers(1, 5))
        cols = sorted(rng.choice(ncols, size=k, replace=False).tolist())
        vals = rng.choice([-2.0, -1.0, 1.0, 2.0, 0.0], size=k, p=[0.1, 0.4, 0.4, 0.05, 0.05])
        rows.append(dict(zip(cols, vals)))
    for i in range(nrows // 4):       # negated and summed rows: dependent
        a, b = rng.integers(len(rows), size=2)
        if rng.random() < 0.5:
            rows.append({j: -v for j, v in rows[a].items()})
        else:
            d = dict(rows[a])
            for j, v in rows[b].items():
                d[j] = d.get(j, 0.0) + v  # may cancel: explicit zero
            rows.append(d)
    order = rng.permutation(len(rows))
    indptr, ind, dat = [0], [], []
    for r in order:
        items = sorted(rows[r].items())
        ind += [j for j, _ in items]
        dat += [v for _, v in items]
        indptr.append(len(ind))
    return sp.csr_matrix((np.array(dat), np.array(ind), np.array(indptr)), shape=(len(rows), ncols))


def test_header_matches_oracle_on_random_constraints(reduce_exe):
    rng = np.random.default_rng(7)
    mats = [_constraint_like(rng, int(rng.integers(3, 60)), int(rng.integers(6, 80))) for _ in range(40)]
    for C, (order, kept) in zip(mats, _run_native(reduce_exe, mats)):
        assert order == colamd(C.shape[1], C.shape[0], C.indptr, C.indices)
        okept, _ = reduce_to_linearly_independent(C)
        assert kept == okept
        assert len(kept) == np.linalg.matrix_rank(C.toarray())


@pytest.mark.parametrize("name", ["icosphere_obj", "ico2", "bumpy3", "bumpy4"])
def test_header_matches_oracle_on_pipeline_constraints(reduce_exe, name):
    """Cfull = constraint * uncompress of the oracle pipeline (explicit zeros included)."""
    from metriko_amd.meshgen import bumpy_icosphere
    from oracle.pipeline import run_pipeline
    from tests.helpers import mesh
    if name.startswith("bumpy"):
        V, F = bumpy_icosphere(int(name[-1]), amp=0.05, jitter=1e-3, seed=5)
    else:
        V, F = mesh(name)
    rp = run_pipeline(V, F, ratio=0.03)["rp"]
    C = structural_product(rp.constraint, rp.uncompress)
    (order, kept), = _run_native(reduce_exe, [C])
    okept, red = reduce_to_linearly_independent(C)
    assert order == colamd(C.shape[1], C.shape[0], C.indptr, C.indices)
    assert kept == okept
    assert red.shape[0] == np.linalg.matrix_rank(C.toarray())
    # the reduced matrix the pipeline went on with is this one, explicit zeros included
    assert np.array_equal(rp.Cfull.indptr, red.indptr) and np.array_equal(rp.Cfull.indices, red.indices)
