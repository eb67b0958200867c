"""Differentiation configuration on the GPU: the sizing hook's sized re-solve and its adjoint
(mk_igm_sized_uv / mk_igm_sized_uv_grad) through the torch binding, against the FP64 CPU
adjoint of oracle/sizing.py and central finite differences.  Bars: with sigma = 1 the sized uv
reproduces the pipeline's uv (1e-8 relative); loss within 1e-8 and gradient within 1e-6
relative (max norm) of the CPU adjoint on the same decisions; the adjoint's dot-product test
within 5e-8 and finite differences of the loss within 1e-4 on the 250k-face configuration."""
import numpy as np
import pytest

import metriko_amd as mk
from metriko_amd.meshgen import genus3_blob
from oracle.pipeline import run_pipeline
from oracle.sizing import loss_and_grad

pytestmark = pytest.mark.gpu


def _u0(nF, seed=4):
    g = np.random.Generator(np.random.PCG64(seed))
    return g.uniform(-1, 1, nF) + 1j * g.uniform(-1, 1, nF)


def _gpu_loss_grad(state, V, F, s):
    import torch
    from metriko_amd.autograd import SizedUV, conformal_distortion, face_scale
    ts = torch.tensor(s, requires_grad=True)
    tF = torch.tensor(F, dtype=torch.long)
    uv = SizedUV.apply(face_scale(ts, tF), state)
    loss = conformal_distortion(uv, torch.tensor(V), tF)
    loss.backward()
    return loss.item(), ts.grad.numpy().copy()


def _run(gpu_ctx, V, F):
    gpu_ctx.keep_igm_state(True)
    try:
        out = gpu_ctx.pipeline(V, F, ratio=0.03, u0=_u0(F.shape[0]))
        state = gpu_ctx.take_igm_state()
    finally:
        gpu_ctx.keep_igm_state(False)
    return out, state


def test_sizing_matches_cpu_adjoint(gpu_ctx):
    V, F = genus3_blob(0.15)
    out, state = _run(gpu_ctx, V, F)
    ref = run_pipeline(V, F, ratio=0.03, u0=_u0(F.shape[0]))
    assert sorted(map(tuple, out["decisions"])) == sorted((i, v) for (i, v, _) in ref["decisions"])
    _, uv1 = state.uv(np.ones(F.shape[0]))
    assert np.abs(uv1 - np.asarray(out["uv"]).reshape(-1)).max() < 1e-8 * np.abs(uv1).max()
    s = 0.1 * np.random.Generator(np.random.PCG64(4)).standard_normal(V.shape[0])
    lg, gg = _gpu_loss_grad(state, V, F, s)
    lc, gc = loss_and_grad(ref["rp"], V, F, s)
    print("loss", lg, lc, "grad rel", np.abs(gg - gc).max() / np.abs(gc).max())
    assert abs(lg - lc) <= 1e-8 * abs(lc)
    assert np.abs(gg - gc).max() <= 1e-6 * np.abs(gc).max()
    state.close()


def test_sizing_250k_finite_differences(gpu_ctx):
    """The configuration's mesh (genus-3 blob, 249,784 faces), log-sizes N(0, 0.1) (seed 4).
    The sized uv is linear in sigma, so the adjoint passes the dot-product test
    g_uv . (uv(sigma + d) - uv(sigma - d)) / 2 = grad(g_uv) . d to solver precision (5e-8);
    the loss gradient in s matches central differences (eps 1e-3) within 1e-4, the precision
    the forward solves' 1e-10 residual leaves to a difference quotient."""
    V, F = genus3_blob()
    out, state = _run(gpu_ctx, V, F)
    print("pipeline", {k: out["info"][k] for k in ("nS", "nT", "k", "success", "ms_total")})
    assert out["info"]["success"] == 1
    g = np.random.Generator(np.random.PCG64(4))
    s = 0.1 * g.standard_normal(V.shape[0])
    sig = np.exp(-(s[F[:, 0]] + s[F[:, 1]] + s[F[:, 2]]) / 3.0)
    guv = g.standard_normal(6 * F.shape[0])
    dsig = 0.1 * g.standard_normal(F.shape[0])
    _, up = state.uv(sig + dsig)
    _, um = state.uv(sig - dsig)
    lhs = guv @ (up - um) / 2.0
    rhs = state.grad(sig, guv) @ dsig
    print("dot-product test", lhs, rhs)
    assert abs(lhs - rhs) <= 5e-8 * abs(lhs)
    loss, gs = _gpu_loss_grad(state, V, F, s)
    for _ in range(2):
        d = g.standard_normal(V.shape[0])
        eps = 1e-3
        lp, _ = _gpu_loss_grad(state, V, F, s + eps * d)
        lm, _ = _gpu_loss_grad(state, V, F, s - eps * d)
        fd = (lp - lm) / (2 * eps)
        print("loss", loss, "fd", fd, "adjoint", gs @ d)
        assert abs(fd - gs @ d) <= 1e-4 * abs(fd)
    state.close()
