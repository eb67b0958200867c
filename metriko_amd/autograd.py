"""Thin torch binding of the sizing hook: per-vertex log-sizes s -> corner uv of the kept
rounding system, differentiable through the library's adjoint solve.

    sigma = face_scale(s, F)                      # exp(-mean log-size per face), torch ops
    uv = SizedUV.apply(sigma, state)             # mk_igm_sized_uv, backward mk_igm_sized_uv_grad
    loss = conformal_distortion(uv, V, F)         # torch ops
    loss.backward()                               # d loss / d s

``state`` is a ``metriko_amd.SizingState`` (``Context.keep_igm_state()`` before
``Context.pipeline``, then ``Context.take_igm_state()``).  FP64 throughout; the library calls
take host arrays, so tensors go through the CPU.
"""
import torch


class SizedUV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma, state):
        _, uv = state.uv(sigma.detach().cpu().double().numpy())
        ctx.state = state
        ctx.save_for_backward(sigma)
        return torch.from_numpy(uv).to(sigma.device)

    @staticmethod
    def backward(ctx, g_uv):
        (sigma,) = ctx.saved_tensors
        g = ctx.state.grad(sigma.detach().cpu().double().numpy(), g_uv.detach().cpu().double().contiguous().numpy())
        return torch.from_numpy(g).to(sigma.device), None


def face_scale(s, F):
    """sigma_f = exp(-(s_a + s_b + s_c) / 3) for per-vertex log-sizes s and faces F (long tensor)."""
    return torch.exp(-(s[F[:, 0]] + s[F[:, 1]] + s[F[:, 2]]) / 3.0)


def conformal_distortion(uv, V, F):
    """sum_f A_f (|J_f|_F^2 - 2 det J_f): J_f maps face f (in its own orthonormal frame) to its
    corner uv (uv[6 f + 2 j + k]); zero exactly where the map is conformal."""
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    n = torch.cross(e1, e2, dim=1)
    A = 0.5 * torch.linalg.norm(n, dim=1)
    x = e1 / torch.linalg.norm(e1, dim=1, keepdim=True)
    y = torch.cross(n / (2 * A)[:, None], x, dim=1)
    Pm = torch.stack([torch.stack([(e1 * x).sum(1), (e2 * x).sum(1)], 1),
                      torch.stack([(e1 * y).sum(1), (e2 * y).sum(1)], 1)], 1)
    U = uv.reshape(-1, 3, 2)
    Q = torch.stack([U[:, 1] - U[:, 0], U[:, 2] - U[:, 0]], dim=2)
    J = Q @ torch.linalg.inv(Pm)
    det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    return torch.sum(A * ((J ** 2).sum(dim=(1, 2)) - 2.0 * det))
