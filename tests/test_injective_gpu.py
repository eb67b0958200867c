"""GPU parity of iterative rounding with localInjectivity = true (mk_iterative_rounding through
the C ABI) against the oracle (oracle/injective.py) on the same inputs.

Cases: the barrier never active (icosphere L2 and a bumpy icosphere L3, grid ratio 0.1), the
barrier active during the rounding (icosphere L2, ratio 0.3), and the naive pass alone
(seamless = false), and a case where the barrier stops the rounding (icosphere L2, ratio 0.6:
a round's LM pass cannot bring its fixed values within 1e-6, success = false).  The decisions
and the success flag agree exactly; fullx to 1e-8 where the rounding succeeds (measured:
1.2e-10 at most; the device solves are iterative to a relative residual of 1e-10).  In the
failing round the barrier derivative reaches 1.7e11 (pairs at r ~ 1e-6 from flipping) and the
damping lambda 1e19, so the damped normal equations are conditioned beyond 1e20: neither a
sparse Cholesky nor PCG fixes those steps to more than a few digits, and fullx of that case
agrees to 1e-2 (measured 1.4e-3).
"""
import numpy as np
import pytest

from metriko_amd.meshgen import icosphere, bumpy_icosphere
from oracle.pipeline import run_pipeline
from tests.helpers import assert_decisions_in_order

pytestmark = pytest.mark.gpu

CASES = {
    "ico2_r0.1": (lambda: icosphere(2), 0.1, True),
    "bumpy3_r0.1": (lambda: bumpy_icosphere(3), 0.1, True),
    "ico2_r0.3_active": (lambda: icosphere(2), 0.3, True),
    "ico2_naive_only": (lambda: icosphere(2), 0.3, False),
    "ico2_r0.6_fails": (lambda: icosphere(2), 0.6, True),
}
TOL = {"ico2_r0.6_fails": 1e-2}


def test_injective_without_callback_matches_oracle(gpu_ctx):
    """iterative_rounding called directly (no iter_cb): the naive pass runs to its own stop."""
    from oracle.injective import run_local_injectivity
    V, F = icosphere(2)
    ref = run_pipeline(V, F, ratio=0.3, local_injectivity=True)
    ir = ref["rp"].ir
    ok_ref = run_local_injectivity(ir)
    fullx, ok, dec, st = gpu_ctx.iterative_rounding(
        ir.fixedIdcs, ir.fixedVals, ir.singularIdcs, ir.integerIdcs, ir.length, ir.Cfull, ir.G2, ir.N, ir.n,
        ir.nF, ir.F2, seamless=True, round_seams=False, local_injectivity=True)
    print(f"[injective] no callback: max |dx| {np.abs(fullx - ir.fullx).max():.3e}", flush=True)
    assert ok == ok_ref
    assert_decisions_in_order(dec, ir.decisions)
    np.testing.assert_allclose(fullx, ir.fullx, atol=1e-8, rtol=0)


@pytest.mark.parametrize("case", list(CASES))
def test_injective_rounding_matches_oracle(gpu_ctx, case):
    mesh, ratio, seamless = CASES[case]
    V, F = mesh()
    ref = run_pipeline(V, F, ratio=ratio, seamless=seamless, local_injectivity=True)
    rp = ref["rp"]
    ir = rp.ir
    print(f"[injective] case {case}: k={ir.k} faces={ir.nF}", flush=True)
    # with RosyParameterization's locally_injectivity_check as the naive pass's iter_cb, as
    # the oracle pipeline runs it
    fullx, ok, dec, st = gpu_ctx.iterative_rounding(
        ir.fixedIdcs, ir.fixedVals, ir.singularIdcs, ir.integerIdcs, ir.length, ir.Cfull, ir.G2, ir.N, ir.n,
        ir.nF, ir.F2, seamless=seamless, round_seams=False, local_injectivity=True,
        cut_idx=np.asarray(rp.cut.idx), fullx2Nfunc=rp.fullx2Nfunc)
    print(f"[injective] case {case}: success {ok} (oracle {ref['success']}), "
          f"max |dx| {np.abs(fullx - ref['fullx']).max():.3e}", flush=True)
    assert ok == ref["success"]
    if seamless:
        assert_decisions_in_order(dec, ref["decisions"])
    np.testing.assert_allclose(fullx, ref["fullx"], atol=TOL.get(case, 1e-8), rtol=0)


def test_gpu_uv_feeds_qex(gpu_ctx):
    """example_qex's flow with the GPU rounding: injective uv from mk_iterative_rounding, corner
    uv as RosyParameterization::integ lays it out, then the host QEx extraction; the same quad
    mesh as the oracle's QEx on the oracle's uv."""
    import metriko_amd as mk
    from oracle import qex
    V, F = icosphere(2)
    ref = run_pipeline(V, F, ratio=0.1, local_injectivity=True)
    rp = ref["rp"]
    ir = rp.ir
    fullx, ok, dec, st = gpu_ctx.iterative_rounding(
        ir.fixedIdcs, ir.fixedVals, ir.singularIdcs, ir.integerIdcs, ir.length, ir.Cfull, ir.G2, ir.N, ir.n,
        ir.nF, ir.F2, seamless=True, round_seams=False, local_injectivity=True,
        cut_idx=np.asarray(rp.cut.idx), fullx2Nfunc=rp.fullx2Nfunc)
    assert ok
    NF = (rp.fullx2Nfunc @ fullx).reshape(-1, ir.N)
    cut = np.asarray(rp.cut.idx)
    uv = np.zeros((3 * len(F), 2))
    for j in range(3):
        uv[j::3] = NF[cut[:, j], :2]
    np.testing.assert_allclose(uv, ref["uv"], atol=1e-8, rtol=0)
    mt = np.asarray(ref["combed_matching"], np.int32)
    sg = np.asarray(ref["singular"], np.int32)
    got = mk.qex_extract(V, F, mt, sg, uv)
    want = qex.extract(ref["mesh"], ref["uv"][:, 0] + 1j * ref["uv"][:, 1], mt, sg)
    print(f"[injective] GPU uv -> QEx: {len(got['faces'])} quads (oracle {len(want['faces'])})", flush=True)
    assert np.array_equal(got["edges"], np.array(want["edges"], np.int32).reshape(-1, 2))
    assert np.array_equal(got["faces"], np.array(want["faces"], np.int32).reshape(-1, 4))
    assert len(got["q_sid"]) - len(got["edges"]) + len(got["faces"]) == 2


def test_gpu_uv_feeds_qgp(gpu_ctx):
    """example_qgp's flow with the GPU rounding: the naive injective pass (seamless = false) on
    the GPU, corner uv, then the host motorcycle graph, T-mesh and quantization; the same
    T-mesh and integer lengths as the oracle's chain on the oracle's uv."""
    import metriko_amd as mk
    from oracle.motorcycle import MotorcycleGraph
    from oracle.quantization import Quantization
    from oracle.tmesh import Tmesh
    V, F = icosphere(2)
    ref = run_pipeline(V, F, ratio=0.1, seamless=False, local_injectivity=True)
    rp = ref["rp"]
    ir = rp.ir
    fullx, ok, dec, st = gpu_ctx.iterative_rounding(
        ir.fixedIdcs, ir.fixedVals, ir.singularIdcs, ir.integerIdcs, ir.length, ir.Cfull, ir.G2, ir.N, ir.n,
        ir.nF, ir.F2, seamless=False, round_seams=False, local_injectivity=True,
        cut_idx=np.asarray(rp.cut.idx), fullx2Nfunc=rp.fullx2Nfunc)
    assert ok
    NF = (rp.fullx2Nfunc @ fullx).reshape(-1, ir.N)
    cut = np.asarray(rp.cut.idx)
    uv = np.zeros((3 * len(F), 2))
    for j in range(3):
        uv[j::3] = NF[cut[:, j], :2]
    np.testing.assert_allclose(uv, ref["uv"], atol=1e-8, rtol=0)
    mt = np.asarray(ref["combed_matching"], np.int32)
    sg = np.asarray(ref["singular"], np.int32)
    got_t = mk.tmesh_build(V, F, mt, sg, uv)
    got_q = mk.qgp_quantize(V, F, mt, sg, uv)
    tm = Tmesh(MotorcycleGraph(ref["mesh"], ref["uv"][:, 0] + 1j * ref["uv"][:, 1], mt, sg))
    want = Quantization(tm).quantize()
    print(f"[injective] GPU uv -> QGP: {len(got_t['tquads'])} T-quads, lengths {got_q['lengths'].astype(int).tolist()[:12]}",
          flush=True)
    assert got_t["tquads"] == [(list(t), list(s)) for t, s in tm.tquads]
    assert got_q["valid"] and np.array_equal(got_q["lengths"], want)
