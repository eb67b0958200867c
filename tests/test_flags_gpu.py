"""GPU parity of iterative_rounding (mk_iterative_rounding through the C ABI) for the flag
combinations the example driver does not use: roundSeams = true (every integer variable is
rounded, in integerIdcs order, iter_rounding_loop.h:118), with and without local
injectivity, and seamless = false without injectivity (the KKT solution x0 is returned,
iter_rounding.h:79-82).  Against the oracle on the same inputs: success and decisions
exact and in order (decisions within 1e-6 of an integer, which need no solve, may
swap inside their run), fullx to 1e-8 (measured differences are below 1e-9)."""
import numpy as np
import pytest

from metriko_amd.meshgen import icosphere, bumpy_icosphere
from oracle.pipeline import run_pipeline
from tests.helpers import assert_decisions_in_order

pytestmark = pytest.mark.gpu

CASES = {
    "ico2_roundseams": (lambda: icosphere(2), 0.1, True, True, False),
    "bumpy3_roundseams": (lambda: bumpy_icosphere(3), 0.1, True, True, False),
    "ico2_roundseams_injective": (lambda: icosphere(2), 0.1, True, True, True),
    "bumpy3_naive_only": (lambda: bumpy_icosphere(3), 0.1, False, False, False),
}


@pytest.mark.parametrize("case", list(CASES))
def test_flags_match_oracle(gpu_ctx, case):
    mesh, ratio, seamless, round_seams, inj = CASES[case]
    V, F = mesh()
    ref = run_pipeline(V, F, ratio=ratio, seamless=seamless, round_seams=round_seams, local_injectivity=inj)
    rp = ref["rp"]
    ir = rp.ir
    print(f"[flags] case {case}: k={ir.k} faces={ir.nF}", flush=True)
    extra = dict(cut_idx=np.asarray(rp.cut.idx), fullx2Nfunc=rp.fullx2Nfunc) if inj else {}
    fullx, ok, dec, st = gpu_ctx.iterative_rounding(
        ir.fixedIdcs, ir.fixedVals, ir.singularIdcs, ir.integerIdcs, ir.length, ir.Cfull, ir.G2, ir.N, ir.n,
        ir.nF, ir.F2, seamless=seamless, round_seams=round_seams, local_injectivity=inj, **extra)
    print(f"[flags] case {case}: success {ok}, {len(dec)} decisions, max |dx| "
          f"{np.abs(fullx - ref['fullx']).max():.3e}", flush=True)
    assert ok == ref["success"]
    assert_decisions_in_order(dec, ref["decisions"])
    np.testing.assert_allclose(fullx, ref["fullx"], atol=1e-8, rtol=0)
