"""Pins the CPU oracle against closed forms, invariants, library routines and brute force.

The reference ships no golden vectors (its only fixture is the input icosphere.obj), so
each oracle function is checked against something that a plausible mistake would break:
glibc's own rand(), exact eigenvectors of planar meshes, Gauss-Bonnet, Poincare-Hopf,
dense eigensolvers, scipy's sparse solvers and the topology of the cut mesh.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.field import (FaceRosyField, GlibcRand, compute_connection, connection_laplacian, curl_matching,
                          eigen_random_complex, field_roots, solve_smallest_eig, compute_singular)
from oracle.hmesh import Hmesh, edge_topology
from oracle.linalg import fullpiv_lu_kernel
from oracle.pipeline import run_pipeline
from oracle.seam import compute_cut_mesh, compute_seam
from tests.helpers import mesh


def planar_grid(n=6, m=5, jitter=0.15, seed=0):
    g = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(n + 1, dtype=float), np.arange(m + 1, dtype=float), indexing="ij")
    V = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], 1)
    inner = (xs.ravel() > 0) & (xs.ravel() < n) & (ys.ravel() > 0) & (ys.ravel() < m)
    V[inner, :2] += jitter * g.uniform(-1, 1, (inner.sum(), 2))
    F = []
    for i in range(n):
        for j in range(m):
            a, b, c, d = i * (m + 1) + j, (i + 1) * (m + 1) + j, (i + 1) * (m + 1) + j + 1, i * (m + 1) + j + 1
            F += [[a, b, c], [a, c, d]]
    return V, np.array(F)


# ---- glibc rand() stream -------------------------------------------------------------------
def test_glibc_rand_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    ref = [libc.rand() for _ in range(2000)]
    g = GlibcRand(1)
    assert [g.next() for _ in range(2000)] == ref
    assert ref[:3] == [1804289383, 846930886, 1681692777]


def test_eigen_random_mapping():
    u = eigen_random_complex(3)
    g = GlibcRand(1)
    r = [g.next() for _ in range(6)]
    assert u[0].real == -1.0 + 2.0 * r[0] / 2147483647.0
    assert u[0].imag == -1.0 + 2.0 * r[1] / 2147483647.0
    assert u[2].imag == -1.0 + 2.0 * r[5] / 2147483647.0


# ---- mesh / geometry -----------------------------------------------------------------------
@pytest.mark.parametrize("name", ["ico2", "icosphere_obj", "torus_small"])
def test_hmesh_invariants(name):
    V, F = mesh(name)
    m = Hmesh(V, F)
    nH = m.nH
    h = np.arange(nH)
    assert np.all(m.twin[m.twin] == h)
    assert np.all(m.next[m.prev] == h)
    assert np.all(m.head == m.tail[m.twin])
    assert np.all(m.edge[m.twin] == m.edge)
    # closed: halfedges 2e / 2e+1, canonical = 2e
    assert np.all(m.edge2half == 2 * np.arange(m.nE))
    # EF(e,0) runs EV0 -> EV1
    for e in range(m.nE):
        f = m.edge2face[e, 0]
        sides = [(F[f, j], F[f, (j + 1) % 3]) for j in range(3)]
        assert (m.edge2vert[e, 0], m.edge2vert[e, 1]) in sides
    # edges sorted by (min, max)
    key = m.edge2vert[:, 0] * (m.nV + 1) + m.edge2vert[:, 1]
    assert np.all(np.diff(key) > 0)
    # vert2half = last halfedge assigned in edge order
    last = {}
    for e in range(m.nE):
        last[m.edge2vert[e, 0]] = 2 * e
        last[m.edge2vert[e, 1]] = 2 * e + 1
    assert all(m.vert2half[v] == last[v] for v in range(m.nV))


@pytest.mark.parametrize("name,chi", [("ico2", 2), ("icosphere_obj", 2), ("torus_small", 0)])
def test_gauss_bonnet_and_angles(name, chi):
    V, F = mesh(name)
    m = Hmesh(V, F)
    assert m.nEularChars == chi
    assert abs(m.angleDefect.sum() - 2 * np.pi * chi) < 1e-10
    # heArgOnFace: 0 on face2half, then pi - interior angle at F1, then + (pi - angle at F2)
    P = m.pos
    for f in range(0, m.nF, 7):
        a, b, c = P[F[f]]

        def ang(p, q, r):
            u, w = q - p, r - p
            return np.arccos(u @ w / np.linalg.norm(u) / np.linalg.norm(w))
        h0, h1, h2 = m.face_halfs(f)
        assert m.heArgOnFace[h0] == 0.0
        assert abs(m.heArgOnFace[h1] - (np.pi - ang(b, c, a))) < 1e-12
        assert abs(m.heArgOnFace[h2] - (2 * np.pi - ang(b, c, a) - ang(c, a, b))) < 1e-12
    if name.startswith("ico"):  # inscribed polyhedron: area just below 4 pi
        assert 0.97 * 4 * np.pi < m.faceArea.sum() < 4 * np.pi


def test_edge_topology_border_orientation():
    V, F = planar_grid(3, 2)
    EV, FE, EF = edge_topology(F)
    for e in range(EV.shape[0]):
        if EF[e, 0] == -1:  # border edge whose only face runs EV1 -> EV0
            f = EF[e, 1]
            sides = [(F[f, j], F[f, (j + 1) % 3]) for j in range(3)]
            assert (EV[e, 1], EV[e, 0]) in sides


# ---- connection Laplacian ------------------------------------------------------------------
@pytest.mark.parametrize("name", ["ico2", "torus_small"])
def test_laplacian_structure(name):
    V, F = mesh(name)
    m = Hmesh(V, F)
    L = connection_laplacian(m, 4)
    assert abs(L - L.getH()).max() < 1e-14
    d = L.diagonal()
    assert np.allclose(d, 3 + 1e-9, atol=1e-15, rtol=0)
    off = (L - sp.diags(d)).tocoo()
    assert np.allclose(np.abs(off.data), 1.0, atol=1e-14)
    assert np.all(np.diff(L.indptr) == 4)


def test_laplacian_planar_closed_form():
    """On a flat patch a constant world direction is parallel: S c = 0 exactly."""
    V, F = planar_grid(6, 5)
    m = Hmesh(V, F)
    N = 4
    L = connection_laplacian(m, N)
    w = np.array([np.cos(0.3), np.sin(0.3), 0.0])
    phi = np.arctan2(m.faceBasisY @ w, m.faceBasisX @ w)
    c = np.exp(1j * N * phi)
    S = L - 1e-9 * sp.identity(m.nF)
    assert np.abs(S @ c).max() < 1e-12
    # and a non-constant direction field is not in the kernel (the check can fail)
    c2 = np.exp(1j * N * (phi + 0.2 * m.baryCenter[:, 0]))
    assert np.abs(S @ c2).max() > 1e-3


# ---- inverse iteration ---------------------------------------------------------------------
def test_smallest_eig_vs_dense_eigh():
    V, F = mesh("ico2")
    m = Hmesh(V, F)
    L = connection_laplacian(m, 4)
    M = m.faceArea
    x = solve_smallest_eig(L, M, eigen_random_complex(m.nF), 50)
    w, U = sla.eigh(L.toarray(), np.diag(M))
    lam = np.vdot(x, L @ x).real / np.vdot(x, M * x).real
    assert abs(np.vdot(x, M * x) - 1.0) < 1e-12  # normalised in the M-norm
    # the iterate lies (to 50-step accuracy) in the lowest eigen-cluster
    assert lam <= w[0] * (1 + 0.05)
    assert np.linalg.norm(L @ x - lam * (M * x)) < 0.05 * np.linalg.norm(L @ x)


def test_smallest_eig_vs_eigsh():
    """Near-degenerate lowest pair on the torus: after 50 steps x lies in the span of the
    lowest eigenvectors (scipy eigsh, shift-invert) with a Rayleigh quotient inside them."""
    V, F = mesh("torus_small")
    m = Hmesh(V, F)
    L = connection_laplacian(m, 4).tocsc()
    M = sp.diags(m.faceArea).tocsc()
    x = solve_smallest_eig(L, m.faceArea, eigen_random_complex(m.nF), 50)
    vals, vecs = spla.eigsh(L, k=6, M=M, sigma=0, which="LM")
    o = np.argsort(vals)
    vals, vecs = vals[o], vecs[:, o]
    lam = np.vdot(x, L @ x).real / np.vdot(x, M @ x).real
    assert vals[0] - 1e-9 <= lam <= vals[2]
    W = vecs[:, :4]
    coef = np.linalg.solve(W.conj().T @ (M @ W), W.conj().T @ (M @ x))
    assert np.linalg.norm(x - W @ coef) / np.linalg.norm(x) < 1e-4  # higher modes decay as (l1/l5)^50


def test_smallest_eig_vs_dense_power():
    """Brute force: 50 dense LAPACK solves of L x = M u reproduce the iterate."""
    V, F = mesh("torus_small")
    m = Hmesh(V, F)
    L = connection_laplacian(m, 4)
    u = eigen_random_complex(m.nF)
    x = solve_smallest_eig(L, m.faceArea, u, 50)
    Ld = L.toarray()
    y = u.copy()
    for _ in range(50):
        y = np.linalg.solve(Ld, m.faceArea * y)
        y = y / np.sqrt(abs(np.vdot(y, m.faceArea * y)))
    assert np.linalg.norm(x - y) / np.linalg.norm(y) < 1e-10


def test_field_roots_closed_form():
    c = np.exp(1j * np.array([0.1, -2.0, 3.0]))
    f = field_roots(c, 4)
    assert np.allclose(f[:, 0] ** 4, c)
    assert np.allclose(f[:, 1] / f[:, 0], 1j)


# ---- matching / singularities --------------------------------------------------------------
@pytest.mark.parametrize("name,total", [("ico2", 8), ("icosphere_obj", 8), ("torus_small", 0), ("ico3", 8)])
def test_poincare_hopf(name, total):
    V, F = mesh(name)
    m = Hmesh(V, F)
    fld = FaceRosyField(m, 4)
    fld.compute_matching("curl")
    assert fld.singular.sum() == total
    assert fld.singular.dtype.kind == "i"


def test_singularities_rotation_invariant():
    V, F = mesh("ico2")
    R = sla.expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.5], [-0.2, 0.5, 0]]))
    s = []
    for W in (V, V @ R.T):
        m = Hmesh(W, F)
        fld = FaceRosyField(m, 4)
        fld.compute_matching("curl")
        s.append(fld.singular.copy())
    assert np.array_equal(s[0], s[1])


# ---- seam / cut ----------------------------------------------------------------------------
@pytest.mark.parametrize("name,genus", [("ico2", 0), ("icosphere_obj", 0), ("torus_small", 1), ("ico3", 0)])
def test_cut_graph_invariants(name, genus):
    r = run_pipeline(*mesh(name), ratio=0.03)
    m, cut, seam, sing = r["mesh"], r["cut"], r["seam"], r["singular"]
    # cutting opens the surface into one topological disk
    assert cut.nV - cut.nE + cut.nF == 1
    sv = set(np.nonzero(sing)[0])
    seam_verts = set(m.edge2vert[seam].ravel())
    assert sv <= seam_verts or seam.sum() == 0
    ne, nv = seam.sum(), len(seam_verts)
    # seam = tree (genus 0) or tree + 2g independent cycles
    assert ne - nv + 1 == 2 * genus


def test_seam_empty_without_singularities_on_sphere_cut():
    """Genus 0 with no singular vertex: the spanning-tree cut retracts to nothing."""
    V, F = mesh("ico2")
    m = Hmesh(V, F)
    seam = compute_seam(m, np.zeros(m.nV, dtype=int))
    assert seam.sum() == 0


# ---- small dense linear algebra ------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_fullpiv_kernel(seed):
    g = np.random.default_rng(seed)
    r, c, rank = 7, 11, 4
    A = g.integers(-2, 3, (r, rank)) @ g.integers(-2, 3, (rank, c))
    A = A.astype(float)
    K, rk = fullpiv_lu_kernel(A)
    assert rk == np.linalg.matrix_rank(A)
    assert K.shape == (c, c - rk)
    assert np.abs(A @ K).max() < 1e-10
    assert np.linalg.matrix_rank(K) == c - rk
    # identity rows on the free columns
    assert np.sum(np.isclose(np.abs(K).max(axis=1), 1.0) & (np.count_nonzero(K, axis=1) == 1)) >= c - rk


# ---- IGM -------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def obj_run():
    return run_pipeline(*mesh("icosphere_obj"), ratio=0.03)


def test_gradient_patch_test(obj_run):
    """G2 applied to the restriction of a linear function gives its exact face gradient."""
    rp = obj_run["rp"]
    cut, N = rp.cut, rp.N
    a = np.array([0.3, -1.2, 0.7])
    vals = cut.pos @ a
    x = np.zeros(N * cut.nV)
    x[0::N] = vals
    g2 = (rp.G2cut @ x).reshape(cut.nF, 2 * N)[:, :2]
    proj = a - (cut.faceNormal @ a)[:, None] * cut.faceNormal
    exact = np.stack([(proj * cut.faceBasisX).sum(1), (proj * cut.faceBasisY).sum(1)], 1)
    assert np.abs(g2 - exact).max() < 1e-10


def test_kkt_against_spsolve(obj_run):
    ir = obj_run["rp"].ir
    E, C, f = ir.E, ir.constMat, ir.f_rhs
    KKT = sp.bmat([[E, C.T], [C, None]], format="csc")
    sol = spla.spsolve(KKT, np.concatenate([f, np.zeros(C.shape[0])]))
    assert np.abs(sol[:ir.k] - ir.XSmall).max() <= 1e-9 * np.abs(sol[:ir.k]).max()
    assert ir.kkt_residual < 1e-12


def test_constraints_hold(obj_run):
    rp = obj_run["rp"]
    Cx = rp.Cfull @ obj_run["fullx"]
    assert np.abs(Cx).max() < 1e-8
    assert np.abs(rp.Cfull @ obj_run["x0"]).max() < 1e-8


def test_integrality_and_seamlessness(obj_run):
    rp, fullx = obj_run["rp"], obj_run["fullx"]
    for (i, v, _) in obj_run["decisions"]:
        assert abs(fullx[i] - v) < 1e-6
    # across every seam edge the two corner uv's differ by a rotation by k*90 deg plus an integer shift
    m, cut = obj_run["mesh"], obj_run["cut"]
    nfn = rp.nfn
    for e in np.nonzero(obj_run["seam"])[0]:
        f0, f1 = m.edge2face[e]
        a, b = m.edge2vert[e]
        ja0, jb0 = list(m.idx[f0]).index(a), list(m.idx[f0]).index(b)
        ja1, jb1 = list(m.idx[f1]).index(a), list(m.idx[f1]).index(b)
        pa0, pb0 = nfn[cut.idx[f0, ja0], :2], nfn[cut.idx[f0, jb0], :2]
        pa1, pb1 = nfn[cut.idx[f1, ja1], :2], nfn[cut.idx[f1, jb1], :2]
        ok = False
        for k in range(4):
            R = np.array([[np.cos(k * np.pi / 2), -np.sin(k * np.pi / 2)], [np.sin(k * np.pi / 2), np.cos(k * np.pi / 2)]])
            t = pa1 - R @ pa0
            if np.abs((pb1 - R @ pb0) - t).max() < 1e-6 and np.abs(t - np.round(t)).max() < 1e-6:
                ok = True
        assert ok, f"seam edge {e} is not seamless"


def test_literal_normal_equations_agree(obj_run):
    """The stable LM formulation agrees with the reference's explicit J^T J solves."""
    lit = run_pipeline(*mesh("ico2"), ratio=0.03, literal=True)
    stb = run_pipeline(*mesh("ico2"), ratio=0.03)
    assert np.abs(lit["uv"] - stb["uv"]).max() / np.abs(stb["uv"]).max() < 1e-9
    assert sorted(d[:2] for d in lit["decisions"]) == sorted(d[:2] for d in stb["decisions"])


def test_quirk_no_singularities_keeps_x0():
    """roundSeams=false and no inside singularity: the rounding loop never runs (iter_rounding.h:99,111)."""
    from oracle.igm import IterativeRounding
    r = run_pipeline(*mesh("ico2"), ratio=0.03)
    ir = r["rp"].ir
    q = IterativeRounding(ir.fixedIdcs, ir.fixedVals, [], ir.integerIdcs, ir.length, ir.Cfull, ir.G2, 4, 2, ir.nF,
                          True, False, ir.F2)
    assert q.run()
    assert np.array_equal(q.fullx, q.x0)
