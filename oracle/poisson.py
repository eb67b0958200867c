"""Oracle: per-mesh u/v Poisson integration of the field direction (batched configuration).

Test infrastructure only (see ``oracle/__init__.py``).  The batched configuration has no
reference counterpart of its own: per mesh it runs the Smoothest field of
``vectorfield/face_rosy_field.h`` (``oracle/field.py``) and then integrates the field's
first direction g_u (root 0, face basis -> 3D) and its quarter turn g_v = n x g_u as two
target gradient fields in the least-squares sense

    min_u  sum_f A_f |grad u - g_f|^2      <=>      L u = d,

with the piecewise-linear stiffness L_ij = sum_f A_f grad phi_i . grad phi_j and
d_i = sum_f A_f grad phi_i . g_f, vertex 0 pinned to 0 (row and column 0 replaced by the
identity, d_0 = 0).  grad phi_i on face (p0, p1, p2) = n x (p_{i+2} - p_{i+1}) / (2 A).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def hat_gradients(V, F):
    """Per face: unit normal n, area A and the three hat-function gradients (nF x 3 x 3)."""
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    dbl = np.sqrt((nrm * nrm).sum(axis=1))
    n = nrm / dbl[:, None]
    A = 0.5 * dbl
    G = np.stack([np.cross(n, p2 - p1), np.cross(n, p0 - p2), np.cross(n, p1 - p0)], axis=1) / dbl[:, None, None]
    return n, A, G


def stiffness(V, F):
    """L_ij = sum_f A_f grad phi_i . grad phi_j (the cotan Laplacian), CSR."""
    _, A, G = hat_gradients(V, F)
    rows, cols, vals = [], [], []
    for i in range(3):
        for j in range(3):
            rows.append(F[:, i])
            cols.append(F[:, j])
            vals.append(A * (G[:, i] * G[:, j]).sum(axis=1))
    nV = V.shape[0]
    L = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nV, nV))
    L.sum_duplicates()
    return L


def divergence(V, F, g):
    """d_i = sum_f A_f grad phi_i . g_f for a per-face 3D vector field g (nF x 3)."""
    _, A, G = hat_gradients(V, F)
    d = np.zeros(V.shape[0])
    for i in range(3):
        np.add.at(d, F[:, i], A * (G[:, i] * g).sum(axis=1))
    return d


def pin_first(L, d):
    """Vertex 0 pinned to 0: row and column 0 of L replaced by the identity, d_0 = 0."""
    L = L.tolil()
    L[0, :] = 0.0
    L[:, 0] = 0.0
    L[0, 0] = 1.0
    d = d.copy()
    d[0] = 0.0
    return L.tocsr(), d


def field_directions(m, compressed, N=4):
    """g_u = root 0 of the N-RoSy field (principal N-th root of the compressed value) in the
    face basis, g_v = n x g_u; both unit, nF x 3."""
    c = np.exp(1j * np.angle(compressed) / N)
    gu = c.real[:, None] * m.faceBasisX + c.imag[:, None] * m.faceBasisY
    gv = np.cross(m.faceNormal, gu)
    return gu, gv


def poisson_uv(V, F, gu, gv):
    """Solves L u = d(g_u), L v = d(g_v) with vertex 0 pinned (sparse direct); returns u, v."""
    L = stiffness(V, F)
    out = []
    for g in (gu, gv):
        Lp, d = pin_first(L, divergence(V, F, g))
        out.append(spla.spsolve(Lp.tocsc(), d))
    return out[0], out[1]
