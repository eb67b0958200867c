"""iterative_rounding's iter_cb through the C ABI (mk_set_naive_iter_callback): the naive LM
pass calls it after every iteration with fullx = UFull x (NaiveIntegration::post_iteration,
iter_rounding_init.h:28; LMSolver::solve, levenberg_marquardt.h:221-224) and stops when it
returns true.  Against the oracle on the same inputs: the values it sees, how often it is
called, and the pass's stopping point (decisions in order, fullx to 1e-8)."""
import numpy as np
import pytest

from metriko_amd.meshgen import icosphere
from oracle.injective import run_local_injectivity
from oracle.pipeline import run_pipeline
from tests.helpers import assert_decisions_in_order, mesh

pytestmark = pytest.mark.gpu


def _args(ir):
    return (ir.fixedIdcs, ir.fixedVals, ir.singularIdcs, ir.integerIdcs, ir.length, ir.Cfull, ir.G2, ir.N, ir.n,
            ir.nF, ir.F2)


@pytest.mark.parametrize("stop", [False, True])
def test_plain_naive_pass_calls_iter_cb_once_with_x0(gpu_ctx, stop):
    """Without local injectivity the naive pass keeps x at the KKT solution and reaches f* in
    its first iteration; the second finds a zero direction and returns before post_iteration:
    one call, with UFull XSmall, and the same result whatever the callback returns."""
    ref = run_pipeline(*mesh("icosphere_obj"), ratio=0.03)
    ir = ref["rp"].ir
    seen = []

    def cb(x):
        seen.append(x)
        return stop

    fullx, ok, dec, _ = gpu_ctx.iterative_rounding(*_args(ir), iter_cb=cb)
    print(f"[iter_cb] plain pass: {len(seen)} call(s)", flush=True)
    assert len(seen) == 1
    x0 = ir.UFull @ ir.XSmall
    np.testing.assert_allclose(seen[0], x0, atol=1e-8 * max(1.0, np.abs(x0).max()), rtol=0)
    assert ok == ref["success"]
    assert_decisions_in_order(dec, ref["decisions"])
    np.testing.assert_allclose(fullx, ref["fullx"], atol=1e-8, rtol=0)


def test_injective_naive_pass_stops_when_iter_cb_says_so(gpu_ctx):
    """Local injectivity on: a callback that returns true ends the naive pass after its first
    iteration; the oracle with the same callback stops at the same iterate."""
    V, F = icosphere(2)
    ir = run_pipeline(V, F, ratio=0.3, local_injectivity=True)["rp"].ir
    calls_ref = []
    ok_ref = run_local_injectivity(ir, lambda X: calls_ref.append(X) or True)
    calls = []
    fullx, ok, dec, _ = gpu_ctx.iterative_rounding(*_args(ir), local_injectivity=True,
                                                   iter_cb=lambda X: calls.append(X) or True)
    print(f"[iter_cb] injective pass: {len(calls)} call(s), oracle {len(calls_ref)}, "
          f"max |dx| {np.abs(fullx - ir.fullx).max():.3e}", flush=True)
    assert len(calls) == len(calls_ref) == 1
    np.testing.assert_allclose(calls[0], calls_ref[0], atol=1e-8, rtol=0)
    assert ok == ok_ref
    assert_decisions_in_order(dec, ir.decisions)
    np.testing.assert_allclose(fullx, ir.fullx, atol=1e-8, rtol=0)
