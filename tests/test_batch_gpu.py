"""Batched configuration on the GPU (mk_batch_run): many closed meshes of uneven size in one
block-diagonal launch set.  Bars: each mesh's field within 1e-8 of the oracle's 50-step inverse
iteration up to a global unit phase (same u0), and u / v within 1e-8 relative of the oracle's
Poisson integration of the same field (oracle/poisson.py); every mesh's own relative residual
at most 1e-10."""
import numpy as np
import pytest

import metriko_amd as mk
from metriko_amd.meshgen import batch_meshes
from oracle.field import connection_laplacian, solve_smallest_eig
from oracle.hmesh import Hmesh
from oracle.poisson import field_directions, poisson_uv
from tests.helpers import mesh, phase_error

pytestmark = pytest.mark.gpu


def _u0(nF, seed):
    g = np.random.Generator(np.random.PCG64(seed))
    return g.uniform(-1, 1, nF) + 1j * g.uniform(-1, 1, nF)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _check_mesh(V, F, u0, o):
    m = Hmesh(V, F)
    ref = solve_smallest_eig(connection_laplacian(m, 4), m.faceArea, u0, 50)
    ferr = phase_error(o["compressed"], ref)
    gu, gv = field_directions(m, o["compressed"], 4)
    u, v = poisson_uv(V, F, gu, gv)
    return ferr, _rel(o["u"], u), _rel(o["v"], v)


def test_batch_small_matches_oracle(gpu_ctx):
    ms = batch_meshes(count=8, seed=3, max_torus_verts=8000, max_level=4)
    u0s = [_u0(F.shape[0], 100 + i) for i, (V, F) in enumerate(ms)]
    out, st = gpu_ctx.batch(ms, 4, 50, u0s)
    print(st)
    assert st["field_rel_residual"] <= 1e-10 and st["poisson_rel_residual"] <= 1e-10
    for (V, F), u0, o in zip(ms, u0s, out):
        ferr, eu, ev = _check_mesh(V, F, u0, o)
        print(V.shape[0], ferr, eu, ev)
        assert ferr < 1e-8 and eu < 1e-8 and ev < 1e-8


def test_batch_64_meshes(gpu_ctx):
    """The 64-mesh set (1,354,535 vertices, 2,708,954 faces): residual bounds, unit mass norm
    per mesh, and the oracle comparison on its meshes of up to 2,562 vertices."""
    ms = batch_meshes()
    u0s = [_u0(F.shape[0], 100 + i) for i, (V, F) in enumerate(ms)]
    out, st = gpu_ctx.batch(ms, 4, 50, u0s)
    print(st)
    assert st["meshes"] == 64 and st["faces"] == 2708954
    assert st["field_rel_residual"] <= 1e-10 and st["poisson_rel_residual"] <= 1e-10
    # every batch of PCG iterations over all 64 meshes replayed one captured hipGraph
    assert st["graph_replays"] > 0 and st["eager_batches"] == 0
    assert st["tuned_poisson_slice_height"] in (64, 128, 256) and st["tuned_poisson_spmv_us"] > 0.0
    checked = 0
    for (V, F), u0, o in zip(ms, u0s, out):
        m_area = Hmesh(V, F).faceArea if V.shape[0] <= 2562 else None
        assert np.isfinite(o["compressed"]).all() and np.isfinite(o["u"]).all() and np.isfinite(o["v"]).all()
        if m_area is not None:
            assert abs(np.sum(m_area * np.abs(o["compressed"]) ** 2) - 1.0) < 1e-12
            ferr, eu, ev = _check_mesh(V, F, u0, o)
            assert ferr < 1e-8 and eu < 1e-8 and ev < 1e-8
            checked += 1
    assert checked >= 8


def test_autotune_layout_gives_same_spmv(gpu_ctx):
    V, F = mesh("torus_small")
    L = connection_laplacian(Hmesh(V, F), 4)
    best, us = gpu_ctx.autotune_spmv(L)
    print("autotune", best.format, best.slice_height, best.sigma, best.block_size, us)
    assert us > 0.0 and best.slice_height in (64, 128, 256) and best.block_size in (128, 256)
    x = np.random.Generator(np.random.PCG64(1)).standard_normal(L.shape[0]) + 0j
    y0 = gpu_ctx.op_complex(L).apply(x)
    y1 = gpu_ctx.op_complex(L, best).apply(x)
    assert np.abs(y1 - y0).max() <= 1e-14 * np.abs(y0).max()
