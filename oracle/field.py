"""Oracle: Knoeppel smoothest N-RoSy face field (FP64, CPU).

Restates ``vectorfield/face_rosy_field.h`` and ``solver/sparse_linear_solver.h``.
Test infrastructure only; see ``oracle/__init__.py``.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import progress
from .hmesh import TWO_PI


def connection_laplacian(m, N):
    """``FaceRosyField::connectionLaplacian`` (face_rosy_field.h:46-76).

    Per edge: r = -e^{i farg(h2)} / e^{i farg(h1)} on h1 = e.half(), 1/r on its
    twin (:52-60).  Per face f and interior neighbour g across h:
    L(f,f) += 1, L(f,g) = -r^N with r = transport[h.twin()] (:62-71).  Returns
    S + 1e-9 I (:73-75).  Complex Hermitian, nF x nF.
    """
    nF, nH = m.nF, m.nH
    transport = np.zeros(nH, dtype=np.complex128)
    h1 = m.edge2half
    h2 = m.twin[h1]
    v1 = np.exp(1j * m.heArgOnFace[h1])
    v2 = np.exp(1j * m.heArgOnFace[h2])
    r = -v2 / v1
    transport[h1] = r
    transport[h2] = 1.0 / r
    rows, cols, vals = [], [], []
    for k in range(3):
        h = m.face2half.copy()
        for _ in range(k):
            h = m.next[h]
        tw = m.twin[h]
        ok = m.face[tw] >= 0
        f = np.arange(nF)[ok]
        g = m.face[tw[ok]]
        rr = transport[tw[ok]]
        rows += [f, f]
        cols += [f, g]
        vals += [np.ones(len(f), dtype=np.complex128), -(rr ** N)]
    S = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nF, nF))
    S.sum_duplicates()
    return (S + 1e-9 * sp.identity(nF, dtype=np.complex128, format="csr")).tocsr()


def mass_matrix(m):
    """``FaceRosyField::galerkinMassMatrix`` (face_rosy_field.h:78-84): diag(face area)."""
    return m.faceArea.copy()


class GlibcRand:
    """glibc ``rand()`` (TYPE_3 additive feedback generator), default seed 1.

    The reference draws its start vector from unseeded ``std::rand``
    (``VecXc::Random``, sparse_linear_solver.h:32,44).  This build fixes that
    stream to glibc's published generator: r[0]=seed,
    r[i]=16807*r[i-1] mod (2^31-1) for i<31, r[i]=r[i-31] for 31<=i<34,
    r[i]=r[i-31]+r[i-3] mod 2^32 afterwards; the k-th output is r[k+344]>>1.
    """

    def __init__(self, seed=1):
        r = [0] * 344
        r[0] = seed
        for i in range(1, 31):
            hi, lo = divmod(r[i - 1], 127773)
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        for i in range(34, 344):
            r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
        self.r = r

    def next(self):
        r = self.r
        v = (r[-31] + r[-3]) & 0xFFFFFFFF
        r.append(v)
        if len(r) > 4096:
            del r[:-64]
        return v >> 1


RAND_MAX = 2147483647


def eigen_random_complex(n, rng=None):
    """Eigen 3.4 ``VecXc::Random(n)``: per entry re then im, each -1 + 2*rand()/RAND_MAX.

    The draw order inside one complex entry is compiler-dependent in the
    reference (argument evaluation order); this build fixes it to real first.
    """
    rng = GlibcRand(1) if rng is None else rng
    out = np.empty(n, dtype=np.complex128)
    for i in range(n):
        re = -1.0 + (2.0 * float(rng.next())) / float(RAND_MAX)
        im = -1.0 + (2.0 * float(rng.next())) / float(RAND_MAX)
        out[i] = complex(re, im)
    return out


def solve_smallest_eig(L, mass, u0, n_iter=50, return_history=False):
    """``solveSmallestEig(SprsC& L, SprsC& M, int nIter = 50)`` (sparse_linear_solver.h:29-54).

    One sparse direct factorisation (SuperLU here, Eigen SparseLU / CHOLMOD in
    the reference), then nIter times: x = L^{-1}(M u); s = sqrt(|x^H M x|);
    x /= s; u = x.  No convergence test.
    """
    lu = spla.splu(L.tocsc())
    progress.report(f"field: L factorised ({L.shape[0]} rows)")
    u = u0.astype(np.complex128).copy()
    x = u
    hist = []
    for it in range(n_iter):
        if it % 10 == 9:
            progress.report(f"field: inverse iteration {it + 1}/{n_iter}")
        x = lu.solve(mass * u)
        s = np.sqrt(abs(np.vdot(x, mass * x)))
        x = x / s
        u = x
        if return_history:
            hist.append(x.copy())
    return (x, hist) if return_history else x


def principal_curvature_dir(m):
    """``FaceRosyField::principalCurvatureDir`` (face_rosy_field.h:86-98).

    Per face, over its halfedges h (face2half, next, next.next): with l = |h| and
    c = l e^{i farg(h)}, dir += -c^2 / l * darg(h); D = dir / 4.  Complex, face basis.
    """
    D = np.zeros(m.nF, dtype=np.complex128)
    h = m.face2half.copy()
    for _ in range(3):
        vec = m.pos[m.head[h]] - m.pos[m.tail[h]]
        length = np.sqrt((vec * vec).sum(axis=1))
        c = np.exp(1j * m.heArgOnFace[h]) * length
        D = D + (-c * c / length * m.dihedralArg[h])
        h = m.next[h]
    return D * 0.25


def solve_curvature_aligned(L, mass, D, N):
    """``FieldType::CurvatureAligned`` (face_rosy_field.h:25-34) with lambda = 0: D squared for
    N = 4, rhs = M D / sqrt(|D^H M D|), compressed = solveSquare(L, rhs) (sparse direct LU,
    sparse_linear_solver.h:16-27)."""
    assert N in (2, 4)
    if N == 4:
        D = D * D
    s = np.sqrt(abs(np.vdot(D, mass * D)))
    rhs = mass * D / s
    return spla.spsolve(L.tocsc(), rhs)


def field_roots(compressed, N):
    """Field from the compressed representation (face_rosy_field.h:37-43):
    field(f,i) = (e^{i arg c_f})^{1/N} * (e^{2 pi i/N})^i, principal branch."""
    th = np.angle(compressed)
    base = np.exp(1j * th / N)
    rot = np.exp(1j * TWO_PI / N * np.arange(N))
    return base[:, None] * rot[None, :]


def compute_connection(m):
    """``computeConnection`` (face_rosy_field.h:109-117): per interior edge eg/ef in the two face frames."""
    conn = np.zeros(m.nE, dtype=np.complex128)
    e = np.nonzero(~m.edge_is_boundary)[0]
    h = m.edge2half[e]
    v = m.pos[m.head[h]] - m.pos[m.tail[h]]
    v = v / np.sqrt((v * v).sum(axis=1))[:, None]
    f0, f1 = m.edge2face[e, 0], m.edge2face[e, 1]
    ef = (v * m.faceBasisX[f0]).sum(1) + 1j * (v * m.faceBasisY[f0]).sum(1)
    eg = (v * m.faceBasisX[f1]).sum(1) + 1j * (v * m.faceBasisY[f1]).sum(1)
    conn[e] = eg / ef
    return conn


def curl_matching(m, field, conn, N):
    """``cMatching`` (face_rosy_field.h:169-202).

    Per interior edge: iM = argmin_i sum_j (t . (v1 - v0))^2 with strict '<'
    (first minimum wins), effort = arg prod_i v1/(v0 conn).
    """
    matching = -np.ones(m.nE, dtype=np.int64)
    effort = np.zeros(m.nE)
    e = np.nonzero(~m.edge_is_boundary)[0]
    h = m.edge2half[e]
    t = m.pos[m.head[h]] - m.pos[m.tail[h]]
    t = t / np.sqrt((t * t).sum(axis=1))[:, None]
    f0, f1 = m.edge2face[e, 0], m.edge2face[e, 1]
    best = np.full(len(e), 32767000.0)
    iM = np.zeros(len(e), dtype=np.int64)
    for i in range(N):
        curl = np.zeros(len(e))
        for j in range(N):
            c0 = field[f0, j]
            c1 = field[f1, (i + j) % N]
            w0 = c0.real[:, None] * m.faceBasisX[f0] + c0.imag[:, None] * m.faceBasisY[f0]
            w1 = c1.real[:, None] * m.faceBasisX[f1] + c1.imag[:, None] * m.faceBasisY[f1]
            curl = curl + ((t * (w1 - w0)).sum(axis=1)) ** 2
        better = curl < best
        iM[better] = i
        best[better] = curl[better]
    matching[e] = iM
    coef = np.ones(len(e), dtype=np.complex128)
    for i in range(N):
        v0 = field[f0, i]
        v1 = field[f1, (i + iM) % N]
        coef = coef * (v1 / (v0 * conn[e]))
    effort[e] = np.angle(coef)
    return matching, effort


def principal_matching(m, field, conn, N):
    """``pMatching`` (face_rosy_field.h:138-167)."""
    matching = -np.ones(m.nE, dtype=np.int64)
    effort = np.zeros(m.nE)
    for e in np.nonzero(~m.edge_is_boundary)[0]:
        f0, f1 = m.edge2face[e]
        c = conn[e]
        minRot = 1000.0
        offset = 0
        coef = 1.0 + 0j
        t0 = c * field[f0, 0]
        for i in range(N):
            jf, jg = field[f0, i], field[f1, i]
            coef *= jg / (c * jf)
            r = np.angle(jg / t0)
            if abs(r) < abs(minRot):
                offset, minRot = i, r
        effort[e] = np.angle(coef)
        res = 0.0
        for i in range(N):
            res += np.angle(field[f1, (i + offset) % N] / (c * field[f0, i]))
        matching[e] = offset - int(_cround((res - effort[e]) / TWO_PI))
    return matching, effort


def _cround(x):
    """C ``std::round``: halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def compute_singular(m, effort, N):
    """``computeSingular`` (face_rosy_field.h:100-107).

    Boundary vertices stay 1; interior: round((N*defect + sum_h (canonical ? -1 : 1)
    * effort[edge(h)]) / 2pi) over the vertex's outgoing halfedges.
    """
    singular = np.ones(m.nV, dtype=np.int64)
    can = m.edge2half[m.edge] == np.arange(m.nH)
    contrib = np.where(can, -1.0, 1.0) * effort[m.edge]
    s = m.angleDefect * N
    acc = np.zeros(m.nV)
    np.add.at(acc, m.tail, contrib)
    s = s + acc
    inner = ~m.isBV
    singular[inner] = _cround(s[inner] / TWO_PI).astype(np.int64)
    return singular


class FaceRosyField:
    """``FaceRosyField(mesh, N, FieldType::Smoothest)`` + ``computeMatching`` (face_rosy_field.h:17-44,119-135)."""

    def __init__(self, m, N, u0=None, n_iter=50, compute=True, kind="smoothest"):
        self.mesh, self.rosyN = m, N
        if not compute:
            return
        self.L = connection_laplacian(m, N)
        self.M = mass_matrix(m)
        if kind == "curvature":
            self.D = principal_curvature_dir(m)
            self.compressed = solve_curvature_aligned(self.L, self.M, self.D, N)
        else:
            self.u0 = eigen_random_complex(m.nF) if u0 is None else u0
            self.compressed = solve_smallest_eig(self.L, self.M, self.u0, n_iter)
        self.field = field_roots(self.compressed, N)

    def compute_matching(self, kind="curl"):
        self.connection = compute_connection(self.mesh)
        if kind == "curl":
            self.matching, self.effort = curl_matching(self.mesh, self.field, self.connection, self.rosyN)
        else:
            self.matching, self.effort = principal_matching(self.mesh, self.field, self.connection, self.rosyN)
        self.singular = compute_singular(self.mesh, self.effort, self.rosyN)
