"""Oracle: iterative rounding with localInjectivity = true (CPU, FP64).

Test infrastructure only (see ``oracle/__init__.py``).  Restates
``iterative_rounding/injective_barrier.h:10-110`` (InjectiveBarrier), the barrier branch of
``iter_rounding_common.h:60-94`` (compute_jacobian), ``levenberg_marquardt.h:13-232``
(DiagonalDamping + LMSolver::solve with lambda = 0.01) and ``iter_rounding.h:33-113`` with
localInjectivity = true: a damped naive LM over (x, f) with the barrier on f, then the
seamless rounding loop whose LM passes carry the barrier on the field X2F x and restart the
damping at 0.01 every round.

Linear algebra: every LM direction solves the reference's damped normal equations
(J^T J + diag(lambda * colsum(J.^2))) d = -J^T E.  The seamless system holds the 1e10 weighted
constraint rows; it is solved as B + wK^2 U^T U by Sherman-Morrison-Woodbury against
B = J^T J without those rows (plus the whole damping diagonal), which is the same matrix.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class InjectiveBarrier:
    """injective_barrier.h:10-110 (s = 0.5)."""

    def __init__(self, N, nF, F2):
        self.N, self.nF = N, nF
        F = np.asarray(F2).reshape(nF, N, 2)
        nrm = np.sqrt((F ** 2).sum(axis=2))
        self.V = nrm * np.roll(nrm, -1, axis=1)              # V(i, j) = |v_j| |v_{j+1}|
        j = np.arange(N)
        base = 2 * N * np.arange(nF)[:, None]
        self.cols = np.stack([base + 2 * j, base + 2 * j + 1, base + 2 * ((j + 1) % N),
                              base + 2 * ((j + 1) % N) + 1], axis=2).reshape(-1, 4)   # J (N nF x 4)
        self.S = np.zeros((nF * N, 4))

    def update(self, field, updateJ):
        """fBarrier (N nF) and, with updateJ, barDer (N nF) and S (the rows of the image map)."""
        N, nF, s = self.N, self.nF, 0.5
        f = np.asarray(field).reshape(nF, N, 2)
        v0, v1 = f, np.roll(f, -1, axis=1)
        v = self.V
        r = (v0[..., 0] * v1[..., 1] - v0[..., 1] * v1[..., 0]) / v
        q = r / s
        bar = q ** 3 - 3 * q ** 2 + 3 * q
        with np.errstate(divide="ignore", invalid="ignore"):
            bar2 = 1.0 / bar - 1.0
        bar2 = np.where(r <= 0, np.inf, bar2)
        bar2 = np.where(r >= s, 0.0, bar2)
        fB = bar2.reshape(-1)
        if not updateJ:
            return fB, None
        d = 3.0 * (r * r / s ** 3) - 6.0 * (r / s ** 2) + 3.0 / s
        d = np.where(r <= 0, np.inf, d)
        d = np.where(r >= s, 0.0, d)
        self.S = np.stack([v1[..., 1] / v, -v1[..., 0] / v, -v0[..., 1] / v, v0[..., 0] / v], axis=2).reshape(-1, 4)
        with np.errstate(divide="ignore", invalid="ignore"):
            der = (-d / (bar * bar)).reshape(-1)
        der = np.where(np.abs(fB) < 10e-9, 0.0, der)
        der = np.where(fB == np.inf, np.inf, der)
        return fB, der

    def image(self, size):
        """gen_image_filed: rows (face, j), the 4 entries of S at columns J (N nF x size)."""
        rows = np.repeat(np.arange(self.nF * self.N), 4)
        return sp.csr_matrix((self.S.reshape(-1), (rows, self.cols.reshape(-1))), shape=(self.nF * self.N, size))


def _damp_vector(J, lam):
    """DiagonalDamping: lambda * sum_rows J^2 per column."""
    return lam * np.asarray(J.multiply(J).sum(axis=0)).ravel()


def lm_damped(x0, objective, solve, maxIter, funcTol, fooTol, lam0, stats, post=None):
    """LMSolver::solve with DiagonalDamping (levenberg_marquardt.h:13-83,140-231).
    objective(x, needJ) -> (E, J); solve(J, damp, rhs) -> d with (J^T J + diag(damp)) d = rhs;
    post(x) -> True ends the solve after an iteration (ST->post_iteration)."""
    lam = lam0
    prevx = x0.copy()
    E, J = objective(prevx, True)
    damp = _damp_vector(J, lam)
    it = 0
    x = prevx
    while True:
        rhs = -(J.T @ E)
        if (np.max(np.abs(rhs)) if rhs.size else 0.0) < fooTol:
            x = prevx  # the reference only breaks here when verbose
        d = solve(J, damp, rhs)
        stats["solves"] += 1
        if np.linalg.norm(d) < funcTol:
            return prevx, True
        Ep, _ = objective(prevx, False)
        En, _ = objective(prevx + d, False)
        e_prev, e_new = Ep @ Ep, En @ En
        if e_prev > e_new:
            x = prevx + d
            if abs(e_prev - e_new) < funcTol:
                return x, False
        else:
            x = prevx
        E, J = objective(x, True)
        # DiagonalDamping::update: energies at prevx and prevx + d decide lambda; the damping
        # diagonal comes from the Jacobian at the new x
        if e_prev > e_new and e_new != np.inf:
            lam /= 10.0
        else:
            lam *= 10.0
        damp = _damp_vector(J, lam)
        if post is not None and post(x):
            return x, True
        it += 1
        prevx = x
        if not it <= maxIter:
            return x, False


def _direct_solve(J, damp, rhs):
    A = (J.T @ J + sp.diags(damp)).tocsc()
    return spla.splu(A).solve(rhs)


def _woodbury_solve(J, damp, rhs, c0, nC, wK):
    """(J^T J + diag(damp)) d = rhs where rows c0 .. c0 + nC of J are wK * U (the constraint
    rows): B = the normal matrix of the other rows + diag(damp), then
    d = y - Z (I / wK^2 + U Z)^{-1} U y with y = B^-1 rhs, Z = B^-1 U^T."""
    Jb = sp.vstack([J[:c0], J[c0 + nC:]]).tocsr()
    U = (J[c0:c0 + nC] / wK).tocsr()
    lu = spla.splu((Jb.T @ Jb + sp.diags(damp)).tocsc())
    Z = lu.solve(U.T.toarray())
    y = lu.solve(rhs)
    M = np.eye(nC) / wK ** 2 + U @ Z
    return y - Z @ np.linalg.solve(M, U @ y)


def locally_injective(fullx2Nfunc, cut_idx, N, X):
    """locally_injectivity_check (parameterization_integ.ipp:45-63): every cut face keeps a
    non-negative orientation in the (u, v) = functions 0, 1 of fullx2Nfunc X."""
    NF = (fullx2Nfunc @ X).reshape(-1, N)
    a, b, c = NF[cut_idx[:, 0], :2], NF[cut_idx[:, 1], :2], NF[cut_idx[:, 2], :2]
    d1, d2 = b - a, c - a
    return not np.any(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0] < 0)


def run_local_injectivity(ir, check=None):
    """iterative_rounding (iter_rounding.h:33-113) with localInjectivity = true on an oracle
    IterativeRounding ``ir`` (its KKT set-up done); fills ir.fullx / ir.decisions and returns
    success.  check(fullx) is the naive pass's iter_cb (NaiveIntegration::post_iteration,
    iter_rounding_init.h:28): True ends that pass."""
    ir.stats = {"solves": 0, "rounds": 0, "lm_rounds": 0}
    N, nF, k = ir.N, ir.nF, ir.k
    F2 = np.asarray(ir.F2)
    l1 = len(F2)
    UF, G2, length = ir.UFull, ir.G2, ir.length
    barrier = InjectiveBarrier(N, nF, F2)
    nfull = UF.shape[0]
    # ---- naive LM over xf_ = (x, f) (iter_rounding_init.h:30-72, weights 1e4/1e4/1/1e-4) ----
    wI, wK, wC, wB = 10e3, 10e3, 1.0, 1e-4
    fix = ir.fixedIdcs
    UExt = sp.bmat([[UF, None], [None, sp.identity(l1)]], format="csr")
    gInteg = (wI * sp.hstack([-length * G2, sp.identity(l1)]) @ UExt).tocsr()
    gClose = (wC * sp.hstack([sp.csr_matrix((l1, nfull)), sp.identity(l1)]) @ UExt).tocsr()
    sel = sp.csr_matrix((np.ones(len(fix)), (np.arange(len(fix)), fix)), shape=(len(fix), nfull + l1))
    gConst = (wK * sel @ UExt).tocsr()

    def obj_n(xf_, needJ):
        xf = UExt @ xf_
        x, f = xf[:nfull], xf[nfull:]
        fInteg = f - length * (G2 @ x)
        fClose = f - F2
        fConst = x[fix] - ir.fixedVals
        fB, der = barrier.update(f, needJ)
        E = np.concatenate([fInteg * wI, fClose * wC, fConst * wK, fB * wB])
        if not needJ:
            return E, None
        gB = sp.diags(der) @ barrier.image(l1) @ gClose[:, k:] * wI
        gB = sp.hstack([sp.csr_matrix((N * nF, k)), gB]).tocsr()
        return E, sp.vstack([gInteg, gClose, gConst, gB]).tocsr()

    XF = np.concatenate([ir.XSmall, F2])
    post = (lambda xf_: check(UF @ xf_[:k])) if check is not None else None
    xn, _ = lm_damped(XF, obj_n, _direct_solve, 100, 10e-10, 10e-10, 0.01, ir.stats, post)
    ir.x_naive = xn
    ir.fullx = UF @ xn[:k]
    if not ir.seamless:
        ir.decisions = []
        return True
    X_ = xn[:k].copy()
    F2_ = xn[k:].copy()
    # ---- seamless loop (iter_rounding_loop.h:10-125, weights 1 / 1e5 / 0.01, barrier 1e-4) ----
    wI, wK, wC, wB = 1.0, 10e4, 0.01, 1e-4
    X2F = (G2 @ UF * length).tocsr()
    X2F.eliminate_zeros()
    left = list(ir.integerIdcs if ir.roundSeams else ir.singularIdcs)
    roundedSingulars = False
    fixedI, fixedV = [], []
    xCurr, xInit = X_.copy(), X_.copy()
    success, rounded, lm_x = True, False, None
    ir.decisions = []
    while left:
        xPrev = xCurr.copy()
        X = UF @ xCurr
        minDif, minIdx = 3276700.0, -1
        for i, li in enumerate(left):
            dd = abs(X[li] - np.floor(X[li] + 0.5) if X[li] >= 0 else X[li] - np.ceil(X[li] - 0.5))
            if dd < minDif:
                minIdx, minDif = i, dd
        pick = left[minIdx]
        v = X[pick]
        rv = int(np.floor(v + 0.5) if v >= 0 else np.ceil(v - 0.5))
        fixedI.append(pick)
        fixedV.append(rv)
        ir.decisions.append((pick, rv, float(minDif)))
        del left[minIdx]
        if not left and not ir.roundSeams and not roundedSingulars:
            left = list(ir.integerIdcs)
            roundedSingulars = True
        ir.stats["rounds"] += 1
        if not (minDif > 10e-7):
            continue
        rounded = True
        ir.stats["lm_rounds"] += 1
        U = UF[fixedI, :].tocsr()
        r = np.array(fixedV, dtype=np.float64)
        gInt, gCl, gCo = wI * X2F, wC * sp.identity(k, format="csr"), wK * U
        nInt = X2F.shape[0]

        def obj_s(xs, needJ, xPrev=xPrev, U=U, r=r, gCo=gCo):
            field = X2F @ xs
            fB, der = barrier.update(field, needJ)
            E = np.concatenate([wI * (field - F2_), wC * (xs - xPrev), wK * (U @ xs - r), wB * fB])
            if not needJ:
                return E, None
            gB = (sp.diags(der) @ barrier.image(nInt) @ X2F * wB).tocsr()
            return E, sp.vstack([gInt, gCl, gCo, gB]).tocsr()

        nC = U.shape[0]

        def solve_s(J, damp, rhs, nC=nC):
            return _woodbury_solve(J, damp, rhs, nInt + k, nC, wK)

        lm_x, _ = lm_damped(xInit, obj_s, solve_s, 100, 1e-7, 1e-7, 0.01, ir.stats)
        fx = UF @ lm_x
        mx = max(abs(fx[i] - np.round(fx[i])) for i in fixedI)
        xCurr, xInit = lm_x.copy(), lm_x.copy()
        if not (mx <= 10e-7):
            success = False
            break
    ir.fixedI, ir.fixedV = fixedI, fixedV
    ir.fullx = UF @ lm_x if rounded else ir.fullx
    return success
