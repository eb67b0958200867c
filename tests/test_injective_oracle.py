"""Pins for the local-injectivity oracle (oracle/injective.py) on the CPU.

The reference holds no fixtures for this path (example_igm runs with localInjectivity =
false), so the oracle is pinned by what the mathematics fixes: the barrier's closed form
and its derivative against finite differences, the damped LM solver on a problem with a
known minimiser, the Woodbury form of the seamless direction against a direct solve, and
the whole path against the plain (lambda = 0, barrier-free) rounding where the barrier
never activates.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from metriko_amd.meshgen import icosphere, bumpy_icosphere
from oracle import injective as inj
from oracle.pipeline import run_pipeline


def _field_pairs(nF, N, rng, angle_lo, angle_hi):
    """nF faces of N vectors; consecutive vectors turn by an angle in [lo, hi]."""
    F = np.zeros((nF, N, 2))
    for i in range(nF):
        a = rng.uniform(0, 2 * np.pi)
        for j in range(N):
            r = rng.uniform(0.5, 2.0)
            F[i, j] = r * np.cos(a), r * np.sin(a)
            a += rng.uniform(angle_lo, angle_hi)
    return F.reshape(-1)


def test_barrier_closed_form():
    N, nF = 4, 1
    ref = np.array([1.0, 0, 0, 1, -1, 0, 0, -1])     # unit 4-RoSy: r = 1 for every pair
    b = inj.InjectiveBarrier(N, nF, ref)
    fB, der = b.update(ref, True)
    assert np.all(fB == 0) and np.all(der == 0)
    # turn vector 1 so that pair (0, 1) has r = s / 2 = 0.25: bar = 1/8 - 3/4 + 3/2 = 7/8
    t = np.arcsin(0.25)
    f = ref.copy()
    f[2:4] = np.cos(t), np.sin(t)
    fB, _ = b.update(f, False)
    assert fB[0] == pytest.approx(1 / 0.875 - 1, rel=1e-14)
    # a flipped pair (r <= 0) is infinite, r >= s is zero
    f[2:4] = np.cos(-0.1), np.sin(-0.1)
    fB, der = b.update(f, True)
    assert fB[0] == np.inf and der[0] == np.inf
    assert fB[2] == 0.0 and fB[3] == 0.0


def test_barrier_jacobian_matches_finite_differences():
    rng = np.random.default_rng(7)
    N, nF = 4, 5
    ref = _field_pairs(nF, N, rng, np.pi / 2, np.pi / 2)
    f = _field_pairs(nF, N, rng, 0.15, 0.45)          # r in (0, s) for most pairs
    b = inj.InjectiveBarrier(N, nF, ref)
    fB, der = b.update(f, True)
    Jb = (sp.diags(der) @ b.image(f.size)).toarray()
    eps = 1e-7
    Jfd = np.zeros_like(Jb)
    for c in range(f.size):
        fp, fm = f.copy(), f.copy()
        fp[c] += eps
        fm[c] -= eps
        Jfd[:, c] = (b.update(fp, False)[0] - b.update(fm, False)[0]) / (2 * eps)
    active = np.isfinite(fB) & (np.abs(fB) > 1e-6)
    assert active.sum() >= 4
    np.testing.assert_allclose(Jb[active], Jfd[active], rtol=1e-5, atol=1e-6)


def test_damped_lm_reaches_known_minimum():
    # Rosenbrock residuals: minimiser (1, 1), energy 0
    def obj(x, needJ):
        E = np.array([10 * (x[1] - x[0] ** 2), 1 - x[0]])
        J = sp.csr_matrix(np.array([[-20 * x[0], 10.0], [-1.0, 0.0]])) if needJ else None
        return E, J
    stats = {"solves": 0}
    x, _ = inj.lm_damped(np.array([-1.2, 1.0]), obj, inj._direct_solve, 100, 1e-12, 1e-12, 0.01, stats)
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-6)
    # lambda = 0 on an affine residual: one Gauss-Newton step lands on the least-squares solution
    A = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 7.0]])
    y = np.array([1.0, -2.0, 0.5])
    lin = lambda x, needJ: (A @ x - y, sp.csr_matrix(A) if needJ else None)
    stats = {"solves": 0}
    x, _ = inj.lm_damped(np.zeros(2), lin, inj._direct_solve, 100, 1e-12, 1e-12, 0.0, stats)
    np.testing.assert_allclose(x, np.linalg.lstsq(A, y, rcond=None)[0], atol=1e-12)


def test_woodbury_direction_matches_direct_solve():
    rng = np.random.default_rng(3)
    k, nInt, nC = 30, 45, 4
    X = sp.random(nInt, k, density=0.2, random_state=3, format="csr")
    U = sp.random(nC, k, density=0.3, random_state=4, format="csr")
    Bar = sp.random(8, k, density=0.3, random_state=5, format="csr")
    J = sp.vstack([X, 0.01 * sp.identity(k), 1e5 * U, Bar]).tocsr()
    damp = 0.01 * np.asarray(J.multiply(J).sum(axis=0)).ravel()
    rhs = rng.standard_normal(k)
    d_direct = inj._direct_solve(J, damp, rhs)
    d_wood = inj._woodbury_solve(J, damp, rhs, nInt + k, nC, 1e5)
    np.testing.assert_allclose(d_wood, d_direct, rtol=1e-7, atol=1e-9 * np.abs(d_direct).max())


@pytest.mark.parametrize("mesh", ["ico2", "bumpy3"])
def test_inactive_barrier_reproduces_plain_rounding(mesh):
    V, F = icosphere(2) if mesh == "ico2" else bumpy_icosphere(3)
    a = run_pipeline(V, F, ratio=0.1, local_injectivity=True)
    b = run_pipeline(V, F, ratio=0.1, local_injectivity=False)
    assert a["success"] and b["success"]
    # the same (index, value) decisions; rounds whose margins tie at round-off may swap order
    assert sorted(d[:2] for d in a["decisions"]) == sorted(d[:2] for d in b["decisions"])
    np.testing.assert_allclose(a["fullx"], b["fullx"], atol=1e-6)


def test_active_barrier_case():
    """icosphere L2, grid ratio 0.3: the barrier activates (pairs under 30 degrees) and the
    rounding still succeeds; the damped path ends within 1e-6 of the plain one."""
    V, F = icosphere(2)
    seen = []
    orig = inj.InjectiveBarrier.update

    def spy(self, field, updateJ):
        fB, der = orig(self, field, updateJ)
        seen.append(int(np.count_nonzero(fB)))
        return fB, der
    inj.InjectiveBarrier.update = spy
    try:
        a = run_pipeline(V, F, ratio=0.3, local_injectivity=True)
    finally:
        inj.InjectiveBarrier.update = orig
    b = run_pipeline(V, F, ratio=0.3, local_injectivity=False)
    assert max(seen) > 0
    assert a["success"]
    np.testing.assert_allclose(a["fullx"], b["fullx"], atol=1e-6)


def test_local_injectivity_check():
    """locally_injectivity_check on two triangles with fullx2Nfunc = identity (N = 2): the
    identity map is injective; moving one corner across the shared edge flips a face."""
    idx = np.array([[0, 1, 2], [0, 2, 3]])
    X = np.array([0.0, 0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0])
    f2n = sp.identity(8, format="csr")
    assert inj.locally_injective(f2n, idx, 2, X)
    Y = X.copy()
    Y[6:8] = 2.0, 0.5          # vertex 3 past the diagonal 0-2: face (0, 2, 3) turns clockwise
    assert not inj.locally_injective(f2n, idx, 2, Y)
    Z = X.copy()
    Z[6:8] = 2.0, 2.0          # vertex 3 on the line 0-2: zero area counts as injective
    assert inj.locally_injective(f2n, idx, 2, Z)
