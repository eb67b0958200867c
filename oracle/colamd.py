"""Oracle: COLAMD column ordering and Eigen's SparseQR rank-revealing column choice.

Test infrastructure only; see ``oracle/__init__.py``.

``reduce_to_linearly_independent`` (solver/matrix_ops.h:12-28) factorises C^T with
``Eigen::SparseQR<SprsD, Eigen::COLAMDOrdering<int>>`` and keeps the rows of C named by the
first ``rank`` entries of ``colsPermutation().indices()``, renumbered in that order.  Both
pieces live in Eigen 3.4 (the version libigl v2.5.0 brings, absent here):

* ``COLAMDOrdering`` runs COLAMD (Davis, Gilbert, Larimore, Ng, "Algorithm 836: COLAMD, a
  column approximate minimum degree ordering algorithm", ACM TOMS 30(3), 2004; Eigen's
  ``Eigen_Colamd.h`` port of COLAMD 2.x) on the pattern of C^T with the default knobs (dense
  rows: more than 0.5 n_col entries; dense columns: more than 0.5 n_row entries; aggressive
  row absorption), and ``colsPermutation`` starts as the inverse of its output.
* ``SparseQR::factorize`` is a left-looking Householder QR over the columns in that order; a
  column whose Householder norm |beta| (its distance to the span of the columns accepted
  before it) is below 20 (m + n) max_j ||A(:,j)|| eps is moved behind the accepted ones, in
  the order of rejection.  So the kept rows of C are the greedy independent subset of the
  COLAMD order, in that order.

``colamd`` below restates the published algorithm step by step (init_rows_cols,
init_scoring, find_ordering with approximate external degrees, mass elimination, aggressive
absorption and hash-based super-column detection, order_children); the "shared" unions of
its Col / Row records are single arrays here, with the same overlapping uses.  Garbage
collection only compacts storage (dead entries are skipped either way), so the column store
simply grows.  Pinned by ``tests/test_colamd.py``: against SuperLU's own COLAMD (scipy's
``splu`` column permutation, which is COLAMD followed by a postorder of the column
elimination tree) on random sparse matrices, and by hand-worked orderings.
"""
import numpy as np

EMPTY = -1
DEAD_PRINCIPAL = -1
DEAD_NON_PRINCIPAL = -2


def colamd(n_row, n_col, colptr, rowind, dense_row=0.5, dense_col=0.5, dense_rule="fraction"):
    """Returns p with p[k] = the column placed k-th (colamd's output array).

    colptr / rowind: the pattern in compressed-column form (row indices sorted, no
    duplicates: Eigen hands over a compressed SparseMatrix).  dense_rule "fraction" is
    Eigen 3.4's (a row is dense above dense_row * n_col entries, a column above
    dense_col * n_row); "sqrt" is the rule of later COLAMD releases (above
    max(16, dense_row sqrt(n_col)) and max(16, dense_col sqrt(min(n_row, n_col))), knobs 10),
    used only to check the restatement against SuperLU's copy of COLAMD.
    """
    colptr = [int(x) for x in colptr]
    nnz = colptr[n_col]
    A = [int(x) for x in rowind[:nnz]] + [0] * nnz   # column form, then row form
    # --- Col / Row records (Colamd_Col, Colamd_Row) ---
    c_start = [0] * n_col
    c_len = [0] * n_col
    c_sh1 = [0] * n_col   # thickness (alive) / parent (dead)
    c_sh2 = [0] * n_col   # score (alive) / order (dead)
    c_sh3 = [0] * n_col   # headhash / hash / prev
    c_sh4 = [0] * n_col   # degree_next / hash_next
    r_start = [0] * n_row
    r_len = [0] * n_row
    r_sh1 = [0] * n_row   # degree / p
    r_sh2 = [0] * n_row   # mark / first_column

    # === init_rows_cols (input sorted, no duplicates: never jumbled) ===
    for c in range(n_col):
        c_start[c] = colptr[c]
        c_len[c] = colptr[c + 1] - colptr[c]
        c_sh1[c] = 1
        c_sh2[c] = 0
        c_sh3[c] = EMPTY
        c_sh4[c] = EMPTY
    for r in range(n_row):
        r_len[r] = 0
        r_sh2[r] = -1
    for c in range(n_col):
        for k in range(colptr[c], colptr[c + 1]):
            r = A[k]
            if r_sh2[r] != c:
                r_len[r] += 1
            else:
                c_len[c] -= 1
            r_sh2[r] = c
    if n_row > 0:
        r_start[0] = nnz
        r_sh1[0] = nnz
        r_sh2[0] = -1
        for r in range(1, n_row):
            r_start[r] = r_start[r - 1] + r_len[r - 1]
            r_sh1[r] = r_start[r]
            r_sh2[r] = -1
    for c in range(n_col):
        for k in range(colptr[c], colptr[c + 1]):
            r = A[k]
            A[r_sh1[r]] = c
            r_sh1[r] += 1
    for r in range(n_row):
        r_sh2[r] = 0
        r_sh1[r] = r_len[r]

    def col_alive(c):
        return c_start[c] >= 0

    def row_alive(r):
        return r_sh2[r] >= 0

    # === init_scoring ===
    if dense_rule == "fraction":
        dense_row_count = max(0, min(int(dense_row * n_col), n_col))
        dense_col_count = max(0, min(int(dense_col * n_row), n_row))
    else:
        dense_row_count = int(max(16.0, dense_row * np.sqrt(n_col)))
        dense_col_count = int(max(16.0, dense_col * np.sqrt(min(n_row, n_col))))
    max_deg = 0
    n_col2 = n_col
    for c in range(n_col - 1, -1, -1):          # empty columns: ordered last, natural order
        if c_len[c] == 0:
            n_col2 -= 1
            c_sh2[c] = n_col2
            c_start[c] = DEAD_PRINCIPAL
    for c in range(n_col - 1, -1, -1):          # dense columns: ordered last, natural order
        if not col_alive(c):
            continue
        if c_len[c] > dense_col_count:
            n_col2 -= 1
            c_sh2[c] = n_col2
            for k in range(c_start[c], c_start[c] + c_len[c]):
                r_sh1[A[k]] -= 1
            c_start[c] = DEAD_PRINCIPAL
    for r in range(n_row):                       # dense and empty rows
        deg = r_sh1[r]
        if deg > dense_row_count or deg == 0:
            r_sh2[r] = -1
        else:
            max_deg = max(max_deg, deg)
    for c in range(n_col - 1, -1, -1):          # initial scores, columns pruned of dead rows
        if not col_alive(c):
            continue
        score = 0
        w = c_start[c]
        for k in range(c_start[c], c_start[c] + c_len[c]):
            r = A[k]
            if not row_alive(r):
                continue
            A[w] = r
            w += 1
            score += r_sh1[r] - 1
            score = min(score, n_col)
        col_length = w - c_start[c]
        if col_length == 0:
            n_col2 -= 1
            c_sh2[c] = n_col2
            c_start[c] = DEAD_PRINCIPAL
        else:
            c_len[c] = col_length
            c_sh2[c] = score
    head = [EMPTY] * (n_col + 1)
    min_score = n_col
    for c in range(n_col - 1, -1, -1):          # degree lists, low indices in front
        if col_alive(c):
            score = c_sh2[c]
            nxt = head[score]
            c_sh3[c] = EMPTY
            c_sh4[c] = nxt
            if nxt != EMPTY:
                c_sh3[nxt] = c
            head[score] = c
            min_score = min(min_score, score)

    # === find_ordering ===
    pfree = 2 * nnz
    tag_mark = 1
    for r in range(n_row):
        if row_alive(r):
            r_sh2[r] = 0
    min_score = 0
    k = 0
    while k < n_col2:
        while head[min_score] == EMPTY and min_score < n_col:
            min_score += 1
        pivot_col = head[min_score]
        nxt = c_sh4[pivot_col]
        head[min_score] = nxt
        if nxt != EMPTY:
            c_sh3[nxt] = EMPTY
        pivot_col_score = c_sh2[pivot_col]
        c_sh2[pivot_col] = k
        pivot_col_thickness = c_sh1[pivot_col]
        k += pivot_col_thickness
        # (garbage collection when pfree + min(pivot_col_score, n_col - k) >= Alen: the store grows instead)
        del pivot_col_score

        # --- pivot row pattern: union of the live rows of the pivot column ---
        pivot_row_start = pfree
        pivot_row_degree = 0
        c_sh1[pivot_col] = -pivot_col_thickness
        for kk in range(c_start[pivot_col], c_start[pivot_col] + c_len[pivot_col]):
            row = A[kk]
            if row_alive(row):
                for rk in range(r_start[row], r_start[row] + r_len[row]):
                    col = A[rk]
                    th = c_sh1[col]
                    if th > 0 and col_alive(col):
                        c_sh1[col] = -th
                        if pfree < len(A):
                            A[pfree] = col
                        else:
                            A.append(col)
                        pfree += 1
                        pivot_row_degree += th
        c_sh1[pivot_col] = pivot_col_thickness
        max_deg = max(max_deg, pivot_row_degree)

        # --- kill every row of the pivot column (the pivot row too, for now) ---
        for kk in range(c_start[pivot_col], c_start[pivot_col] + c_len[pivot_col]):
            r_sh2[A[kk]] = -1

        pivot_row_length = pfree - pivot_row_start
        pivot_row = A[c_start[pivot_col]] if pivot_row_length > 0 else EMPTY

        # --- set differences |row \ pivot row| of the rows met in the pivot row's columns ---
        for rp in range(pivot_row_start, pivot_row_start + pivot_row_length):
            col = A[rp]
            col_thickness = -c_sh1[col]
            c_sh1[col] = col_thickness
            cur_score = c_sh2[col]
            prev_col = c_sh3[col]
            next_col = c_sh4[col]
            if prev_col == EMPTY:
                head[cur_score] = next_col
            else:
                c_sh4[prev_col] = next_col
            if next_col != EMPTY:
                c_sh3[next_col] = prev_col
            for kk in range(c_start[col], c_start[col] + c_len[col]):
                row = A[kk]
                row_mark = r_sh2[row]
                if row_mark < 0:
                    continue
                set_difference = row_mark - tag_mark
                if set_difference < 0:
                    set_difference = r_sh1[row]
                set_difference -= col_thickness
                if set_difference == 0:        # aggressive absorption
                    r_sh2[row] = -1
                else:
                    r_sh2[row] = set_difference + tag_mark

        # --- add up set differences per column, compact, mass elimination, hash ---
        for rp in range(pivot_row_start, pivot_row_start + pivot_row_length):
            col = A[rp]
            hsh = 0
            cur_score = 0
            w = c_start[col]
            for kk in range(c_start[col], c_start[col] + c_len[col]):
                row = A[kk]
                row_mark = r_sh2[row]
                if row_mark < 0:
                    continue
                A[w] = row
                w += 1
                hsh += row
                cur_score += row_mark - tag_mark
                cur_score = min(cur_score, n_col)
            c_len[col] = w - c_start[col]
            if c_len[col] == 0:
                c_start[col] = DEAD_PRINCIPAL
                pivot_row_degree -= c_sh1[col]
                c_sh2[col] = k
                k += c_sh1[col]
            else:
                c_sh2[col] = cur_score
                hsh %= n_col + 1
                head_column = head[hsh]
                if head_column > EMPTY:
                    first_col = c_sh3[head_column]
                    c_sh3[head_column] = col
                else:
                    first_col = -(head_column + 2)
                    head[hsh] = -(col + 2)
                c_sh4[col] = first_col
                c_sh3[col] = hsh

        # --- super-column detection (detect_super_cols) ---
        for rp in range(pivot_row_start, pivot_row_start + pivot_row_length):
            col = A[rp]
            if not col_alive(col):
                continue
            hsh = c_sh3[col]
            head_column = head[hsh]
            if head_column > EMPTY:
                first_col = c_sh3[head_column]
            else:
                first_col = -(head_column + 2)
            super_c = first_col
            while super_c != EMPTY:
                length = c_len[super_c]
                prev_c = super_c
                c = c_sh4[super_c]
                while c != EMPTY:
                    if c_len[c] != length or c_sh2[c] != c_sh2[super_c]:
                        prev_c = c
                        c = c_sh4[c]
                        continue
                    s1, s2 = c_start[super_c], c_start[c]
                    same = all(A[s1 + i] == A[s2 + i] for i in range(length))
                    if not same:
                        prev_c = c
                        c = c_sh4[c]
                        continue
                    c_sh1[super_c] += c_sh1[c]
                    c_sh1[c] = super_c
                    c_start[c] = DEAD_NON_PRINCIPAL
                    c_sh2[c] = EMPTY
                    c_sh4[prev_c] = c_sh4[c]
                    c = c_sh4[c]
                super_c = c_sh4[super_c]
            if head_column > EMPTY:
                c_sh3[head_column] = EMPTY
            else:
                head[hsh] = EMPTY

        c_start[pivot_col] = DEAD_PRINCIPAL
        tag_mark += max_deg + 1            # marks below tag_mark read as "not seen"

        # --- finalize the pivot row and the column scores ---
        w = pivot_row_start
        for rp in range(pivot_row_start, pivot_row_start + pivot_row_length):
            col = A[rp]
            if not col_alive(col):
                continue
            A[w] = col
            w += 1
            A[c_start[col] + c_len[col]] = pivot_row
            c_len[col] += 1
            cur_score = c_sh2[col] + pivot_row_degree
            max_score = n_col - k - c_sh1[col]
            cur_score -= c_sh1[col]
            cur_score = min(cur_score, max_score)
            c_sh2[col] = cur_score
            nxt = head[cur_score]
            c_sh4[col] = nxt
            c_sh3[col] = EMPTY
            if nxt != EMPTY:
                c_sh3[nxt] = col
            head[cur_score] = col
            min_score = min(min_score, cur_score)
        if pivot_row_degree > 0:
            r_start[pivot_row] = pivot_row_start
            r_len[pivot_row] = w - pivot_row_start
            r_sh1[pivot_row] = pivot_row_degree
            r_sh2[pivot_row] = 0

    # === order_children ===
    for i in range(n_col):
        if c_start[i] != DEAD_PRINCIPAL and c_sh2[i] == EMPTY:
            parent = i
            while True:
                parent = c_sh1[parent]
                if c_start[parent] == DEAD_PRINCIPAL:
                    break
            c = i
            order = c_sh2[parent]
            while True:
                c_sh2[c] = order
                order += 1
                c_sh1[c] = parent
                c = c_sh1[c]
                if c_sh2[c] != EMPTY:
                    break
            c_sh2[parent] = order
    p = [0] * n_col
    for c in range(n_col):
        p[c_sh2[c]] = c
    return p


def sparse_qr_kept_columns(M, p):
    """Eigen ``SparseQR::factorize`` rank decisions on the dense matrix M (m x n) whose columns
    are visited in the order p: returns the accepted columns, in order.

    A column is accepted when the Householder norm |beta| of its part outside the span of the
    accepted columns reaches 20 (m + n) max_j ||M(:, j)|| eps (the default threshold; a zero
    matrix uses 1 for the norm)."""
    M = np.asarray(M, dtype=np.float64)
    m, n = M.shape
    if n == 0:
        return []
    max2 = float(np.max(np.sqrt((M * M).sum(axis=0)))) if m else 0.0
    if max2 == 0.0:
        max2 = 1.0
    thr = 20.0 * (m + n) * max2 * np.finfo(np.float64).eps
    diag = min(m, n)
    Hs = []          # Householder vectors (full length m) and their taus
    kept = []
    for j in p:
        v = M[:, j].copy()
        for (hv, tau, piv) in Hs:   # apply H_piv = I - tau hv hv^T in acceptance order
            v -= tau * (hv @ v) * hv
        nz = len(kept)
        if nz >= diag:
            continue
        c0 = v[nz]
        sq = float(v[nz + 1:] @ v[nz + 1:])
        if sq == 0.0:
            beta = c0
            tau = 0.0
            hv = np.zeros(m)
            hv[nz] = 1.0
        else:
            beta = np.sqrt(c0 * c0 + sq)
            if c0 >= 0.0:
                beta = -beta
            hv = np.zeros(m)
            hv[nz] = 1.0
            hv[nz + 1:] = v[nz + 1:] / (c0 - beta)
            tau = (beta - c0) / beta
        if abs(beta) >= thr:
            Hs.append((hv, tau, nz))
            kept.append(j)
    return kept


def structural_product(A, B):
    """A*B as Eigen's default (conservative) sparse product stores it: every (i, j) reached by
    some A(i,k) B(k,j) is an entry, explicit zeros included; CSR with sorted indices."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    B = sp.csr_matrix(B)
    Ab = sp.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
    Bb = sp.csr_matrix((np.ones(B.nnz), B.indices, B.indptr), shape=B.shape)
    P = (Ab @ Bb).tocsr()
    P.sort_indices()
    V = (A @ B).tocsr()
    data = np.zeros(P.nnz)
    for i in range(P.shape[0]):
        lo, hi = P.indptr[i], P.indptr[i + 1]
        vlo, vhi = V.indptr[i], V.indptr[i + 1]
        vals = dict(zip(V.indices[vlo:vhi].tolist(), V.data[vlo:vhi].tolist()))
        data[lo:hi] = [vals.get(int(j), 0.0) for j in P.indices[lo:hi]]
    return sp.csr_matrix((data, P.indices.copy(), P.indptr.copy()), shape=P.shape)


def reduce_to_linearly_independent(C):
    """``reduce_to_linearly_independent`` (matrix_ops.h:12-28) on C (CSR, sorted indices; its
    stored entries, explicit zeros included, are the pattern COLAMD sees).

    Returns (kept row indices in their new order, the reduced C with the kept rows' stored
    entries, explicit zeros included)."""
    import scipy.sparse as sp
    C = sp.csr_matrix(C)
    C.sort_indices()
    nr, nc = C.shape
    if nr == 0:
        return [], C
    # C^T in compressed-column form is C in compressed-row form
    p = colamd(nc, nr, C.indptr, C.indices)
    dense = np.zeros((nc, nr))
    for i in range(nr):
        for kk in range(C.indptr[i], C.indptr[i + 1]):
            dense[C.indices[kk], i] = C.data[kk]
    kept = sparse_qr_kept_columns(dense, p)
    indptr = [0]
    cols, vals = [], []
    for r in kept:
        lo, hi = C.indptr[r], C.indptr[r + 1]
        cols.extend(C.indices[lo:hi].tolist())
        vals.extend(C.data[lo:hi].tolist())
        indptr.append(len(cols))
    out = sp.csr_matrix((np.asarray(vals, dtype=np.float64), np.asarray(cols, dtype=np.int64),
                         np.asarray(indptr, dtype=np.int64)), shape=(len(kept), nc))
    return kept, out
