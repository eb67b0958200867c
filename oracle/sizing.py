"""Oracle: the sizing hook of the differentiation configuration (CPU, FP64, SuperLU).

Test infrastructure only (see ``oracle/__init__.py``); no reference counterpart.  The target
field of face f is scaled by sigma_f = exp(-mean of the face's per-vertex log-sizes), and the
iterative rounding's last LM round (``iter_rounding_loop.h:77-86`` with the integer decisions,
the proximal point xPrev and the Woodbury columns of that round held fixed) is re-solved:

    (A0 + wK^2 U^T U) x = wI^2 X2F^T (sigma .* F2) + wC^2 xPrev + wK^2 U^T r,

x = xt + Z M^{-1} (r - U xt), xt = A0^{-1} b, M = U Z + I / wK^2 (X2F carries length here);
sigma = 1 gives that round's x*.  Corner uv = P UFull x (the rows of fullx2Nfunc that
``parameterization_integ.ipp:88-117`` reads per face corner).  The gradient of a scalar of the
uv is the adjoint solve H lambda = UFull^T P^T g_uv, d/d sigma_f = wI^2 (X2F lambda)_f . F2_f.
Loss of the configuration: the area-weighted conformal distortion sum_f A_f (|J_f|_F^2 -
2 det J_f) = sum_f A_f (s1 - s2)^2 of the uv map J_f of each face.
"""
import numpy as np
import scipy.sparse as sp


def corner_map(rp):
    """P (6 nF x nfull): uv[6 f + 2 j + k] = row N cut.idx(f, j) + k of fullx2Nfunc."""
    N = rp.N
    A = rp.fullx2Nfunc.tocsr()
    idx = rp.cut.idx
    nF = idx.shape[0]
    rows = (N * idx[:, :, None] + np.arange(2)[None, None, :]).reshape(-1)
    return A[rows, :].tocsr(), nF


def face_scale(s, F):
    """sigma_f = exp(-(s_a + s_b + s_c) / 3)."""
    return np.exp(-(s[F[:, 0]] + s[F[:, 1]] + s[F[:, 2]]) / 3.0)


def sized_uv(rp, sigma):
    ir, L = rp.ir, rp.ir.last_round
    P, nF = corner_map(rp)
    F2 = np.asarray(ir.F2).reshape(nF, -1)
    b = L["wI"] ** 2 * (L["X2F"].T @ (sigma[:, None] * F2).reshape(-1)) + L["wC"] ** 2 * L["xPrev"]
    xt = L["A0lu"].solve(b)
    x = xt + L["Z"] @ np.linalg.solve(L["M"], L["r"] - L["Uf"] @ xt)
    fullx = ir.UFull @ x
    return fullx, P @ fullx


def sized_grad(rp, sigma, g_uv):
    ir, L = rp.ir, rp.ir.last_round
    P, nF = corner_map(rp)
    F2 = np.asarray(ir.F2).reshape(nF, -1)
    g = ir.UFull.T @ (P.T @ g_uv)
    y = L["A0lu"].solve(g)
    lam = y - L["Z"] @ np.linalg.solve(L["M"], L["Uf"] @ y)
    w = (L["X2F"] @ lam).reshape(nF, -1)
    return L["wI"] ** 2 * (w * F2).sum(axis=1)


def conformal_distortion(uv, V, F):
    """sum_f A_f (|J_f|_F^2 - 2 det J_f) and its gradient with respect to uv (6 nF)."""
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    n = np.cross(e1, e2)
    A = 0.5 * np.linalg.norm(n, axis=1)
    x = e1 / np.linalg.norm(e1, axis=1)[:, None]
    y = np.cross(n / (2 * A)[:, None], x)
    Pm = np.stack([np.stack([(e1 * x).sum(1), (e2 * x).sum(1)], 1),
                   np.stack([(e1 * y).sum(1), (e2 * y).sum(1)], 1)], 1)          # nF x 2 x 2
    Pi = np.linalg.inv(Pm)
    U = uv.reshape(-1, 3, 2)
    Q = np.stack([U[:, 1] - U[:, 0], U[:, 2] - U[:, 0]], axis=2)                   # nF x 2 x 2 (rows u, v)
    J = Q @ Pi
    det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    loss = float(np.sum(A * ((J ** 2).sum(axis=(1, 2)) - 2.0 * det)))
    # dL/dJ = A (2 J - 2 cof(J)), cof = [[J11, -J10], [-J01, J00]]
    cof = np.stack([np.stack([J[:, 1, 1], -J[:, 1, 0]], 1), np.stack([-J[:, 0, 1], J[:, 0, 0]], 1)], 1)
    gJ = 2.0 * A[:, None, None] * (J - cof)
    gQ = gJ @ np.transpose(Pi, (0, 2, 1))
    gU = np.zeros_like(U)
    gU[:, 1] += gQ[:, :, 0]
    gU[:, 2] += gQ[:, :, 1]
    gU[:, 0] -= gQ[:, :, 0] + gQ[:, :, 1]
    return loss, gU.reshape(-1)


def loss_and_grad(rp, V, F, s):
    """Loss of the sized uv and its gradient with respect to the per-vertex log-sizes s."""
    sigma = face_scale(s, F)
    _, uv = sized_uv(rp, sigma)
    loss, guv = conformal_distortion(uv, V, F)
    gsig = sized_grad(rp, sigma, guv)
    gs = np.zeros(V.shape[0])
    c = -gsig * sigma / 3.0
    for j in range(3):
        np.add.at(gs, F[:, j], c)
    return loss, gs
