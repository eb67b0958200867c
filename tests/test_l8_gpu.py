"""Configuration 5 at its full size on the GPU: the icosphere of level 8 (655,362 vertices,
1,310,720 faces) with radial displacement 0.05 sin(6 theta) sin(5 phi).

* The benchmark's workload (radial jitter N(0, 6.25e-5), the level-4 jitter scaled with the
  edge length): the whole field + IGM pipeline in FP32-multigrid + FP64 (MK_PREC_MIXED_MG)
  against plain FP64 PCG, both to a 1e-10 relative residual.  Integers equal (singularities,
  matching, seam, cut-mesh indices, the set of rounding decisions; their order may differ
  only between candidates whose distances to the grid agree to round-off), field within 1e-8
  up to a global phase, uv within 1e-8 relative.
* The configuration's own jitter N(0, 1e-3): the field stage (the IGM of that mesh, with
  16.9k singularities and 87.6k constraint rows, is not run): both precisions to a 1e-10
  relative residual and phase error between them within 1e-8; the singularity counts are
  reported (on this folded surface some vertices' index sums sit at a rounding tie, so fields
  equal to 1e-15 can differ in a few singularities).
The oracle does not run at this size (minutes of SuperLU); the sizes it does run are covered
by tests/test_gpu_parity.py and tests/test_golden.py."""
import numpy as np
import pytest

import metriko_amd as mk
from tests.helpers import phase_error

pytestmark = pytest.mark.gpu


def _workload(jitter=None):
    from bench import workload
    return workload(8, jitter=jitter)


def test_l8_pipeline_mixed_matches_fp64(gpu_ctx):
    V, F, u0 = _workload()
    print(f"[L8] bench workload: {V.shape[0]} vertices, {F.shape[0]} faces", flush=True)
    out = {}
    for name, prec in (("mixed", mk.PREC_MIXED_MG), ("fp64", mk.PREC_FP64)):
        o = mk.opts(precision=prec)
        out[name] = gpu_ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=o, igm_opts=o)
        inf = out[name]["info"]
        print(f"[L8] {name}: nS {inf['nS']} nT {inf['nT']} field residual {inf['field']['rel_residual']:.2e} "
              f"IGM residual {inf['igm']['solve']['rel_residual']:.2e} KKT {inf['igm']['kkt_rel_residual']:.2e} "
              f"success {inf['success']}", flush=True)
        assert inf["field"]["rel_residual"] <= 1e-10
        assert inf["igm"]["solve"]["rel_residual"] <= 1e-10
        assert inf["success"] == 1
    a, b = out["mixed"], out["fp64"]
    for key in ("singular", "matching", "seam", "cut_idx"):
        assert np.array_equal(a[key], b[key]), key
    assert (a["info"]["nS"], a["info"]["nT"], a["info"]["k"]) == (b["info"]["nS"], b["info"]["nT"], b["info"]["k"])
    assert sorted(a["decisions"]) == sorted(b["decisions"])
    err = phase_error(a["compressed"], b["compressed"])
    uv = np.abs(a["uv"] - b["uv"]).max() / np.abs(b["uv"]).max()
    print(f"[L8] mixed vs fp64: field phase {err:.2e}, uv {uv:.2e}, decisions in the same order: "
          f"{a['decisions'] == b['decisions']}", flush=True)
    assert err <= 1e-8
    assert uv <= 1e-8


def test_l8_as_specified_field_stage(gpu_ctx):
    V, F, u0 = _workload(jitter=1e-3)
    he = mk.hmesh_halfedges(V, F)
    x = {}
    for name, prec in (("mixed", mk.PREC_MIXED_MG), ("fp64", mk.PREC_FP64)):
        L = gpu_ctx.field_laplacian(he["face_halfs"], he["twin"], he["half_face"], he["farg"], 4)
        x[name], st = gpu_ctx.smallest_eig(L, he["face_area"], u0, 50, mk.opts(precision=prec))
        L.close()
        print(f"[L8 jitter 1e-3] {name}: {st['iterations']} iterations, residual {st['rel_residual']:.2e}",
              flush=True)
        assert st["rel_residual"] <= 1e-10
    err = phase_error(x["mixed"], x["fp64"])
    print(f"[L8 jitter 1e-3] phase error {err:.2e}", flush=True)
    assert err <= 1e-8
    # the folded surface (jitter a quarter of an edge) puts some vertices' index sums at a
    # rounding tie: fields equal to 1e-15 can then differ in a few singularities, so the
    # integers are reported here, not compared
    ca, cb = mk.combinatorics(V, F, x["mixed"]), mk.combinatorics(V, F, x["fp64"])
    print(f"[L8 jitter 1e-3] nS {ca['nS']} / {cb['nS']}, nT {ca['nT']} / {cb['nT']}, constraint rows "
          f"{ca['constraint_rows']} / {cb['constraint_rows']}, singular vertices that differ "
          f"{int((ca['singular'] != cb['singular']).sum())}", flush=True)
