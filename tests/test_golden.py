"""Regression vectors on the reference's own input fixture (example/models/icosphere.obj),
written by tests/golden/make_golden.py from the pinned oracle."""
import os

import numpy as np
import pytest

import metriko_amd as mk
from oracle.field import eigen_random_complex
from oracle.pipeline import run_pipeline
from tests.helpers import GOLDEN, assert_decisions_in_order, mesh, phase_error


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "icosphere_obj_oracle.npz"), allow_pickle=False)


def test_oracle_reproduces_golden(golden):
    V, F = mesh("icosphere_obj")
    r = run_pipeline(V, F, ratio=0.03, u0=eigen_random_complex(F.shape[0]))
    assert phase_error(r["compressed"], golden["compressed"]) < 1e-12
    assert np.array_equal(r["singular"], golden["singular"])
    assert np.array_equal(np.nonzero(r["seam"])[0], golden["seam_edges"])
    assert np.array_equal(r["cut"].idx, golden["cut_idx"])
    assert np.array_equal(np.array([(i, v) for (i, v, _) in r["decisions"]]), golden["decisions"])
    assert np.abs(r["uv"] - golden["uv"]).max() < 1e-9


def test_host_stages_reproduce_golden(golden):
    V, F = mesh("icosphere_obj")
    c = mk.combinatorics(V, F, golden["compressed"])
    assert np.array_equal(c["singular"], golden["singular"])
    assert np.array_equal(np.nonzero(c["seam"])[0], golden["seam_edges"])
    assert np.array_equal(c["cut_idx"], golden["cut_idx"])
    assert np.array_equal(c["matching"], golden["combed_matching"])
    assert (c["nS"], c["nT"]) == (int(golden["nS"]), int(golden["nT"]))


@pytest.mark.gpu
def test_gpu_pipeline_reproduces_golden(gpu_ctx, golden):
    V, F = mesh("icosphere_obj")
    out = gpu_ctx.pipeline(V, F, ratio=0.03, u0=eigen_random_complex(F.shape[0]))
    assert phase_error(out["compressed"], golden["compressed"]) < 1e-8
    assert np.array_equal(out["singular"], golden["singular"])
    assert np.array_equal(np.nonzero(out["seam"])[0], golden["seam_edges"])
    ref = [(i, v, m) for (i, v), m in zip(golden["decisions"].tolist(), golden["decision_margin"].tolist())]
    assert_decisions_in_order(out["decisions"], ref)
    assert np.abs(out["uv"] - golden["uv"]).max() / np.abs(golden["uv"]).max() < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [mk.PREC_MIXED_MG, mk.PREC_FP64])
def test_gpu_pipeline_torus_204800(gpu_ctx, precision):
    """The 204,800-face torus (R = 1, r = 0.35, 256 x 400, normal jitter 0.002) against the
    oracle's fixture (tests/golden/make_golden.py --torus): singularities, seam, cut-mesh
    indices (checksum) and rounding decisions bit-exact, the field within 1e-8 up to a global
    phase and uv within 1e-8 relative on sampled faces / corners, in both precision modes."""
    from tests.helpers import cut_hash, torus_config
    g = np.load(os.path.join(GOLDEN, "torus_256x400_oracle.npz"), allow_pickle=False)
    V, F, u0 = torus_config()
    o = mk.opts(precision=precision)
    out = gpu_ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=o, igm_opts=o)
    print(precision, out["info"])
    assert phase_error(out["compressed"][g["field_faces"]], g["field"]) < 1e-8
    sing = np.nonzero(out["singular"])[0]
    assert np.array_equal(sing, g["singular_v"]) and np.array_equal(out["singular"][sing], g["singular_idx"])
    assert np.array_equal(np.nonzero(out["seam"])[0], g["seam_edges"])
    assert out["info"]["cutV"] == int(g["cutV"]) and cut_hash(out["cut_idx"]) == int(g["cut_hash"])
    assert (out["info"]["nS"], out["info"]["nT"]) == (int(g["nS"]), int(g["nT"]))
    ref = [(i, v, m) for (i, v), m in zip(g["decisions"].tolist(), g["decision_margin"].tolist())]
    assert_decisions_in_order(out["decisions"], ref)
    uv = np.asarray(out["uv"]).reshape(-1, 2)[g["uv_corners"]]
    assert np.abs(uv - g["uv"]).max() / np.abs(g["uv"]).max() < 1e-8


@pytest.mark.gpu
def test_gpu_pipeline_l8_matches_oracle_fixture(gpu_ctx):
    """The benchmark's 1,310,720-face mesh (bench.workload(8)) against the oracle's fixture
    (tests/golden/make_golden.py --l8; 12 minutes of single-thread SuperLU on the CPU):
    singularities, seam, cut-mesh indices (checksum), nS / nT and the 74 rounding decisions
    bit-exact, the field within 1e-8 up to a global phase on every 500th face and uv within
    1e-8 relative on every 1500th corner, in the default precision (FP32 multigrid inside FP64
    PCG)."""
    from bench import workload
    from tests.helpers import cut_hash
    g = np.load(os.path.join(GOLDEN, "bumpy_L8_oracle.npz"), allow_pickle=False)
    V, F, u0 = workload(8)
    o = mk.opts(precision=mk.PREC_MIXED_MG)
    out = gpu_ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=o, igm_opts=o)
    err = phase_error(out["compressed"][g["field_faces"]], g["field"])
    uv = np.asarray(out["uv"]).reshape(-1, 2)[g["uv_corners"]]
    uv_err = np.abs(uv - g["uv"]).max() / np.abs(g["uv"]).max()
    print(f"[L8 oracle] field phase {err:.2e}, uv {uv_err:.2e}, decisions {len(out['decisions'])}", flush=True)
    assert err < 1e-8
    sing = np.nonzero(out["singular"])[0]
    assert np.array_equal(sing, g["singular_v"]) and np.array_equal(out["singular"][sing], g["singular_idx"])
    assert np.array_equal(np.nonzero(out["seam"])[0], g["seam_edges"])
    assert out["info"]["cutV"] == int(g["cutV"]) and cut_hash(out["cut_idx"]) == int(g["cut_hash"])
    assert (out["info"]["nS"], out["info"]["nT"]) == (int(g["nS"]), int(g["nT"]))
    assert out["info"]["success"] == int(g["success"])
    ref = [(i, v, m) for (i, v), m in zip(g["decisions"].tolist(), g["decision_margin"].tolist())]
    assert_decisions_in_order(out["decisions"], ref)
    assert uv_err < 1e-8

