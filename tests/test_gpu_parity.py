"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same inputs.

Bars: integers (singularities, matching, seam, cut-mesh indices, rounding decisions)
bit-exact; field within 1e-8 of the oracle up to one global unit phase; uv within 1e-8
relative (max abs difference / max |uv|).  Every solve runs to a 1e-10 FP64 relative
residual (the library default).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import metriko_amd as mk
from oracle.field import connection_laplacian, eigen_random_complex, solve_smallest_eig
from oracle.hmesh import Hmesh
from oracle.pipeline import run_pipeline
from tests.helpers import assert_decisions_in_order, mesh, phase_error

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-8
UV_TOL = 1e-8


@pytest.mark.parametrize("name", ["ico2", "icosphere_obj", "torus_small"])
def test_field_laplacian_device_assembly(gpu_ctx, name):
    V, F = mesh(name)
    he = mk.hmesh_halfedges(V, F)
    Lop = gpu_ctx.field_laplacian(he["face_halfs"], he["twin"], he["half_face"], he["farg"], 4)
    Lg = Lop.export_csr()
    Lo = connection_laplacian(Hmesh(V, F), 4)
    d = (Lg - Lo).tocoo()
    assert Lg.nnz == Lo.nnz
    assert np.abs(d.data).max() < 1e-13
    # Hermitian, |offdiag| = 1
    assert abs(Lg - Lg.getH()).max() < 1e-14


@pytest.mark.parametrize("precision", [mk.PREC_FP64, mk.PREC_MIXED, mk.PREC_MIXED_MG])
@pytest.mark.parametrize("name", ["ico2", "icosphere_obj", "torus_small"])
def test_smallest_eig_matches_oracle(gpu_ctx, name, precision):
    V, F = mesh(name)
    m = Hmesh(V, F)
    he = mk.hmesh_halfedges(V, F)
    Lop = gpu_ctx.field_laplacian(he["face_halfs"], he["twin"], he["half_face"], he["farg"], 4)
    u0 = eigen_random_complex(m.nF)
    x, st = gpu_ctx.smallest_eig(Lop, he["face_area"], u0, 50, mk.opts(precision=precision))
    ref = solve_smallest_eig(connection_laplacian(m, 4), m.faceArea, u0, 50)
    err = phase_error(x, ref)
    print(name, precision, "field err", err, st)
    assert err < FIELD_TOL
    assert st["rel_residual"] <= 1e-10


@pytest.mark.parametrize("precision", [mk.PREC_FP64, mk.PREC_MIXED, mk.PREC_MIXED_MG])
@pytest.mark.parametrize("name", ["ico2", "icosphere_obj", "torus_small"])
def test_smallest_eig_matches_oracle(gpu_ctx, name, precision):
    V, F = mesh(name)
    m = Hmesh(V, F)
    he = mk.hmesh_halfedges(V, F)
    Lop = gpu_ctx.field_laplacian(he["face_halfs"], he["twin"], he["half_face"], he["farg"], 4)
    u0 = eigen_random_complex(m.nF)
    x, st = gpu_ctx.smallest_eig(Lop, he["face_area"], u0, 50, mk.opts(precision=precision))
    ref = solve_smallest_eig(connection_laplacian(m, 4), m.faceArea, u0, 50)
    err = phase_error(x, ref)
    print(name, precision, "field err", err, st)
    assert err < FIELD_TOL
    assert st["rel_residual"] <= 1e-10


def _ordered_real_decisions(dec_gpu, dec_ref):
    """Decisions whose oracle margin exceeds 1e-6 trigger a solve; their order must agree.
    Decisions within 1e-6 of an integer (no solve follows) may tie at round-off level."""
    real = {(i, v) for (i, v, m) in dec_ref if m > 1e-6}
    a = [d for d in dec_gpu if tuple(d) in real]
    b = [(i, v) for (i, v, m) in dec_ref if m > 1e-6]
    return a, b


@pytest.mark.parametrize("precision", [mk.PREC_FP64, mk.PREC_MIXED, mk.PREC_MIXED_MG])
@pytest.mark.parametrize("name", ["ico2", "icosphere_obj", "torus_small"])
def test_pipeline_matches_oracle(gpu_ctx, name, precision):
    V, F = mesh(name)
    nF = F.shape[0]
    u0 = eigen_random_complex(nF)
    ref = run_pipeline(V, F, ratio=0.03, u0=u0)
    o = mk.opts(precision=precision)
    out = gpu_ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=o, igm_opts=o)
    info = out["info"]
    print(name, precision, info)
    assert phase_error(out["compressed"], ref["compressed"]) < FIELD_TOL
    assert np.array_equal(out["singular"], ref["singular"])
    assert np.array_equal(out["seam"], ref["seam"])
    assert np.array_equal(out["matching"], ref["combed_matching"])
    assert np.array_equal(out["cut_idx"], ref["cut"].idx)
    assert info["nS"] == ref["nS"] and info["nT"] == ref["nT"]
    assert_decisions_in_order(out["decisions"], ref["decisions"])
    assert info["success"] == int(ref["success"])
    uv_err = np.abs(out["uv"] - ref["uv"]).max() / np.abs(ref["uv"]).max()
    print("uv rel err", uv_err)
    assert uv_err < UV_TOL


def test_pipeline_is_deterministic(gpu_ctx):
    V, F = mesh("icosphere_obj")
    u0 = eigen_random_complex(F.shape[0])
    a = gpu_ctx.pipeline(V, F, ratio=0.03, u0=u0)
    b = gpu_ctx.pipeline(V, F, ratio=0.03, u0=u0)
    assert np.array_equal(a["uv"], b["uv"])
    assert np.array_equal(a["compressed"], b["compressed"])


def test_spd_solve_real_and_breakdown(gpu_ctx):
    n = 500
    rng = np.random.default_rng(0)
    A = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n) + 0.01, -np.ones(n - 1)], [-1, 0, 1], format="csr")
    b = rng.standard_normal(n)
    op = gpu_ctx.op_real(A)
    for prec in (mk.PREC_FP64, mk.PREC_MIXED, mk.PREC_MIXED_MG):
        x, st = op.solve(b, o=mk.opts(precision=prec))
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10
    # negated matrix: CG breakdown (Jacobi) and the multigrid setup both report MK_ERR_NOT_SPD
    opn = gpu_ctx.op_real(-A)
    for pc in (mk.PRECOND_JACOBI, mk.PRECOND_AMG):
        with pytest.raises(mk.MetrikoError) as ei:
            opn.solve(b, o=mk.opts(precision=mk.PREC_FP64, precond=pc))
        assert ei.value.status == 3


@pytest.mark.gpu
@pytest.mark.parametrize("m", [5, 37, 70])
def test_spd_solve_many_matches_direct(gpu_ctx, m):
    """Batched multi-right-hand-side PCG (blocks of up to 64 columns) against SuperLU, with the
    sequential FP64 path as a second check; every column to a 1e-10 relative residual."""
    nx = 70
    T = sp.diags([-np.ones(nx - 1), 2.0 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(nx)) + sp.kron(sp.identity(nx), T) + 1e-3 * sp.identity(nx * nx)).tocsr()
    rng = np.random.default_rng(m)
    B = rng.standard_normal((A.shape[0], m))
    B[:, 1] = 0.0  # an empty column converges at once
    op = gpu_ctx.op_real(A)
    Xref = spla.splu(A.tocsc()).solve(B)
    for prec in (mk.PREC_MIXED_MG, mk.PREC_FP64):
        X, st = op.solve_many(B, o=mk.opts(precision=prec))
        print(m, prec, st)
        assert st["solves"] == m
        res = np.linalg.norm(A @ X - B, axis=0) / np.maximum(np.linalg.norm(B, axis=0), 1e-300)
        assert res.max() <= 1e-10
        assert np.abs(X[:, 1]).max() == 0.0
        err = np.linalg.norm(X - Xref, axis=0) / np.maximum(np.linalg.norm(Xref, axis=0), 1e-300)
        assert err.max() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mk.PREC_FP64, mk.PREC_MIXED, mk.PREC_MIXED_MG])
def test_graph_replay_matches_eager(gpu_ctx, precision):
    """Multigrid PCG iterations replayed from captured hipGraphs give bit-identical results to
    individual launches, replay after replay, with eager solves in between."""
    nx = 60
    T = sp.diags([-np.ones(nx - 1), 2.0 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(nx)) + sp.kron(sp.identity(nx), T) + 1e-3 * sp.identity(nx * nx)).tocsr()
    op = gpu_ctx.op_real(A)
    o = mk.opts(precision=precision)
    rng = np.random.default_rng(precision)
    rhs = [rng.standard_normal(A.shape[0]) for _ in range(3)]
    try:
        gpu_ctx.set_graphs(False)
        eager = [op.solve(b, o=o)[0] for b in rhs]
        for rep in range(3):
            gpu_ctx.set_graphs(True)
            for b, xe in zip(rhs, eager):
                xg, st = op.solve(b, o=o)
                assert np.array_equal(xg, xe), (rep, st)
            gpu_ctx.set_graphs(False)
            assert np.array_equal(op.solve(rhs[0], o=o)[0], eager[0])
    finally:
        gpu_ctx.set_graphs(True)
    assert np.linalg.norm(A @ eager[0] - rhs[0]) / np.linalg.norm(rhs[0]) <= 1e-10


@pytest.mark.gpu
def test_kernel_timing_leaves_results_unchanged(gpu_ctx):
    """Per-kernel timing (start / end stamps carried by the launches) changes no result: a
    multigrid solve and a block solve with timing on equal the same solves with it off, bit for
    bit, and the report holds every timed kernel with its calls, time and bytes."""
    nx = 60
    T = sp.diags([-np.ones(nx - 1), 2.0 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(nx)) + sp.kron(sp.identity(nx), T) + 1e-3 * sp.identity(nx * nx)).tocsr()
    op = gpu_ctx.op_real(A)
    o = mk.opts(precision=mk.PREC_MIXED_MG)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(A.shape[0])
    B = rng.standard_normal((A.shape[0], 40))
    x_off = op.solve(b, o=o)[0]
    X_off = op.solve_many(B, o=o)[0]
    try:
        gpu_ctx.set_timing(True)
        gpu_ctx.timing_report(reset=True)
        x_on = op.solve(b, o=o)[0]
        X_on = op.solve_many(B, o=o)[0]
        rep = gpu_ctx.timing_report(reset=True)
    finally:
        gpu_ctx.set_timing(False)
    assert np.array_equal(x_on, x_off) and np.array_equal(X_on, X_off)
    assert rep and all(k["calls"] > 0 and k["ms"] > 0.0 and k["bytes"] > 0.0 for k in rep)
    assert any(":spmm" in k["key"] for k in rep) and any(":cheb" in k["key"] for k in rep)


@pytest.mark.gpu
def test_operator_rejects_foreign_context(gpu_ctx):
    """Captured graphs are cached on the operator and hold its context's buffers, so an operator
    is only run on the context it was created on: another context gets MK_ERR_ARG, and the
    owning context still solves (with graphs on) afterwards."""
    nx = 40
    T = sp.diags([-np.ones(nx - 1), 2.0 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(nx)) + sp.kron(sp.identity(nx), T) + 1e-3 * sp.identity(nx * nx)).tocsr()
    op = gpu_ctx.op_real(A)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x0, _ = op.solve(b)
    with mk.Context(0) as other:
        foreign = mk.Op(other, op.h, op.n, op.cplx)
        try:
            with pytest.raises(mk.MetrikoError) as ei:
                foreign.solve(b)
            assert ei.value.status == 1
        finally:
            foreign.h = None  # the handle belongs to gpu_ctx's operator
    x1, _ = op.solve(b)
    assert np.array_equal(x0, x1)
