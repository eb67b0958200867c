"""Oracle: Integer-Grid-Map parameterization with iterative rounding (FP64, CPU).

Restates ``igm/parameterization.h``, ``igm/parameterization_setup.ipp``,
``igm/parameterization_integ.ipp``, ``igm/iterative_rounding/iter_rounding*.h``,
``solver/levenberg_marquardt.h`` and ``solver/matrix_ops.h`` for the flag set
of ``example/example_igm.cpp:23-31`` (seamless, no local injectivity).
Test infrastructure only; see ``oracle/__init__.py``.

Numerics.  Every Levenberg-Marquardt pass of the reference runs with lambda = 0
(``iter_rounding.h:178,217,223``) on an affine residual, so each direction it
computes is d = x* - x_prev with x* the exact minimiser of the quadratic
objective.  The reference forms J^T J explicitly and factorises it; with the
1e4..1e5 weights that product has condition numbers of 1e11..1e14 and the
computed d carries round-off of up to ~1e-6 relative in smooth modes.  This
oracle keeps the LM control flow (iteration, acceptance and stopping tests)
but evaluates x* in an exactly equivalent, well-conditioned form:

* naive pass (``iter_rounding_init.h``): the f-block of J^T J is diagonal and is
  eliminated (Schur complement); the penalty on the fixed DOFs only pins the
  translations that B = length*G2*UFull annihilates, so x* equals the KKT
  solution and f* = (wI^2 B x + wC^2 F2) / (wI^2 + wC^2);
* rounding passes (``iter_rounding_loop.h``): A = A0 + wK^2 U^T U with
  A0 = wI^2 X2F^T X2F + wC^2 I fixed for the whole loop, solved as
  x* = xt + Z (wK^-2 I + U Z)^{-1} (r - U xt), xt = A0^{-1} b0, Z = A0^{-1} U^T.

``literal=True`` switches both passes to the reference's own explicit
normal-equation solve, for small meshes, to measure that round-off.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import progress
from .colamd import reduce_to_linearly_independent as _reduce_colamd, structural_product
from .linalg import fullpiv_lu_kernel


def _prune_zeros(A):
    """Eigen ``prune(1e-3)`` / ``pruned()`` on integer-valued entries: drop |x| <= 1e-15."""
    A = A.tocsr().copy()
    A.data[np.abs(A.data) <= 1e-15] = 0.0
    A.eliminate_zeros()
    return A


def _cround(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


class RosyParameterization:
    """``RosyParameterization`` (parameterization.h:11-123, setup.ipp, integ.ipp)."""

    def __init__(self, raw, cut, ext, singular, matching, seam, degree, ratio=0.02):
        self.raw, self.cut, self.ext = raw, cut, ext
        self.singular, self.matching, self.seam = singular, matching, seam
        self.gridscale = ratio
        self.seamless, self.roundSeams, self.verbose, self.localInjectivity = False, True, False, False
        N = degree
        self.N = N
        if N % 2 == 0:  # set_sign_symmetry (parameterization.h:84-90)
            self.n = N // 2
            self.lreductor = np.vstack([np.eye(self.n), -np.eye(self.n)])
        else:
            self.n = N
            self.lreductor = np.eye(N)
        self.fixedVals = np.zeros(self.n)
        self.compute_he2transidx()
        self.compute_he2matching()
        self.nR = raw.nV + self.nT
        self.nS = int(sum(1 for v in range(raw.nV) if self.is_inside_singular(v)))

    def is_inside_singular(self, v):
        return (not self.raw.isBV[v]) and self.singular[v] != 0

    def find_first_seam_he(self, v):
        """parameterization.h:81: first seam halfedge clockwise from v.half(), else v.half()."""
        for h in self.raw.vert_halfs_cw(v):
            if self.seam[self.raw.edge[h]]:
                return h
        return self.raw.vert2half[v]

    def find_first_bndr_he(self, v):
        """parameterization.h:82."""
        for h in self.raw.vert_halfs_cw(v):
            if self.raw.edge_is_boundary[self.raw.edge[h]]:
                return h
        return self.raw.vert2half[v]

    def compute_he2matching(self):
        """setup.ipp:3-9."""
        m = self.raw
        can = m.edge2half[m.edge] == np.arange(m.nH)
        mm = np.where(can, -1, 1) * self.matching[m.edge]
        self.he2matching = mm % self.N

    def compute_he2transidx(self):
        """setup.ipp:11-50: trace seam segments between non-regular seam vertices."""
        m = self.raw
        he = np.full(m.nH, 32767, dtype=np.int64)
        val = np.zeros(m.nV, dtype=np.int64)
        for e in range(m.nE):
            if self.seam[e]:
                val[m.edge2vert[e, 0]] += 1
                val[m.edge2vert[e, 1]] += 1
        claimed = np.zeros(m.nE, dtype=bool)
        tid = 1
        for v in range(m.nV):
            if (val[v] == 2 and not self.is_inside_singular(v)) or val[v] == 0:
                continue
            for cH in m.vert_halfs_ccw(v, self.find_first_bndr_he(v)):
                cE = m.edge[cH]
                if m.edge_is_boundary[cE] or not self.seam[cE] or claimed[cE]:
                    continue
                he[cH] = tid
                he[m.twin[cH]] = -tid
                claimed[cE] = True
                jv = m.head[cH]
                while val[jv] == 2 and not self.is_inside_singular(jv) and not m.isBV[jv]:
                    nh = None
                    for h in m.vert_halfs_ccw(jv):
                        if self.seam[m.edge[h]] and not claimed[m.edge[h]]:
                            nh = h
                            break
                    he[nh] = tid
                    he[m.twin[nh]] = -tid
                    claimed[m.edge[nh]] = True
                    jv = m.head[nh]
                tid += 1
        self.he2transidx = he
        self.nT = tid - 1

    def setup(self):
        """setup.ipp:52-170: transition permutations around every vertex."""
        raw, cut, N, n = self.raw, self.cut, self.N, self.n
        unit = np.zeros((N, N), dtype=np.int64)
        for i in range(N):
            unit[(i + 1) % N, i] = 1
        cpm = [np.eye(N, dtype=np.int64)]
        for i in range(1, N):
            cpm.append(unit @ cpm[-1])
        vT, cT = [], []
        cur = 0
        I = np.eye(N, dtype=np.int64)

        def assign_block(T, mat, r0, c0):
            for i in range(mat.shape[0]):
                for j in range(mat.shape[1]):
                    if mat[i, j] != 0:
                        T.append((r0 + i, c0 + j, float(mat[i, j])))

        for v in range(raw.nV):
            permMats = [I.copy()]
            permIdcs = [v]
            iVcut = -1
            start = self.find_first_bndr_he(v) if raw.isBV[v] else self.find_first_seam_he(v)
            for h in raw.vert_halfs_ccw(v, start):
                if raw.face[h] == -1:
                    break
                f = raw.face[h]
                jVcut = -1
                for j in range(3):
                    if raw.idx[f, j] == v:
                        jVcut = cut.idx[f, j]
                if jVcut != iVcut:
                    iVcut = jVcut
                    for i in range(len(permIdcs)):
                        assign_block(vT, permMats[i], N * iVcut, N * permIdcs[i])
                hn = raw.twin[raw.prev[h]]
                if raw.face[hn] != -1 and self.seam[raw.edge[hn]]:
                    p = cpm[self.he2matching[hn]]
                    t = self.he2transidx[hn]
                    if t > 0:
                        permMats = [p @ mm for mm in permMats]
                        permMats.append(I.copy())
                        permIdcs.append(raw.nV + t - 1)
                    else:
                        permMats.append(-I)
                        permIdcs.append(raw.nV - t - 1)
                        permMats = [p @ mm for mm in permMats]
            if not raw.isBV[v]:
                idcs = sorted(set(permIdcs))
                mats = []
                for j in range(len(idcs)):
                    mj = np.zeros((N, N), dtype=np.int64)
                    for k in range(len(permIdcs)):
                        if idcs[j] == permIdcs[k]:
                            mj = mj + permMats[k]
                    if idcs[j] == v:
                        mj = mj - I
                    mats.append(mj)
                if any(np.abs(mm).max() != 0 for mm in mats):
                    for j in range(len(mats)):
                        assign_block(cT, mats[j], N * cur, N * idcs[j])
                    cur += 1

        def trip(T, shape):
            if not T:
                return sp.csr_matrix(shape)
            r, c, x = zip(*T)
            return sp.csr_matrix((x, (r, c)), shape=shape)

        self.vtrans2cut = _prune_zeros(trip(vT, (N * cut.nV, N * self.nR)))
        self.constraint = _prune_zeros(trip(cT, (N * cur, N * self.nR)))
        buff = []
        for i in range(0, N * self.nR, N):
            assign_block(buff, self.lreductor.astype(np.int64), i, i * n // N)
        self.uncompress = trip(buff, (N * self.nR, n * self.nR))
        if self.nS == 0:
            self.fixedIdcs = np.arange(n)
        else:
            iV = next(v for v in range(raw.nV) if self.is_inside_singular(v))
            self.fixedIdcs = n * iV + np.arange(n)
        self.integerIdcs = np.array([n * (raw.nV + i) + j for i in range(self.nT) for j in range(n)], dtype=np.int64)
        self.singularIdcs = np.array([n * v + j for v in range(raw.nV) if self.is_inside_singular(v)
                                      for j in range(n)], dtype=np.int64)

    def gradient_operators(self):
        """G2 and G3 (integ.ipp:9-28): per cut face and corner, g = n x e_opp / (2A)."""
        cut, N = self.cut, self.N
        nF = cut.nF
        r2, c2, x2, r3, c3, x3 = [], [], [], [], [], []
        for k3 in range(3):
            h = cut.face2half.copy()
            for _ in range(k3):
                h = cut.next[h]
            nh = cut.next[h]
            e = cut.pos[cut.head[nh]] - cut.pos[cut.tail[nh]]
            g = np.cross(cut.faceNormal, e) / (cut.faceArea * 2.0)[:, None]
            gx = (g * cut.faceBasisX).sum(1)
            gy = (g * cut.faceBasisY).sum(1)
            f = np.arange(nF)
            for k in range(N):
                c = N * cut.tail[h] + k
                er = 3 * N * f + 3 * k
                ir = 2 * N * f + 2 * k
                for comp in range(3):
                    r3.append(er + comp); c3.append(c); x3.append(g[:, comp])
                r2 += [ir, ir + 1]; c2 += [c, c]; x2 += [gx, gy]
        G3 = sp.csr_matrix((np.concatenate(x3), (np.concatenate(r3), np.concatenate(c3))),
                           shape=(3 * N * nF, N * cut.nV))
        G2 = sp.csr_matrix((np.concatenate(x2), (np.concatenate(r2), np.concatenate(c2))),
                           shape=(2 * N * nF, N * cut.nV))
        return G2, G3

    def integ(self, literal=False):
        """integ.ipp:3-120 (with the fixed example_igm flags)."""
        raw, cut, N, n = self.raw, self.cut, self.N, self.n
        self.fullx2Nfunc = (self.vtrans2cut @ self.uncompress).tocsr()
        Cfull = structural_product(self.constraint, self.uncompress)
        Cfull = reduce_to_linearly_independent(Cfull)
        self.Cfull = Cfull
        G2, G3 = self.gradient_operators()
        self.G2cut, self.G3 = G2, G3
        ext = self.ext
        norm = np.mean(np.sqrt((ext.reshape(cut.nF, N, 3) ** 2).sum(axis=2)))  # avg_norm (parameterization.h:113-121)
        F2 = np.zeros((cut.nF, 2 * N))
        for i in range(N):
            blk = ext[:, 3 * i:3 * i + 3]
            F2[:, 2 * i] = (blk * cut.faceBasisX).sum(1)
            F2[:, 2 * i + 1] = (blk * cut.faceBasisY).sum(1)
        self.rawF2Vec = F2.reshape(-1)
        bb = cut.pos.max(axis=0) - cut.pos.min(axis=0)
        self.length = np.sqrt(bb @ bb) * self.gridscale / norm
        self.G2full = (G2 @ self.fullx2Nfunc).tocsr()
        progress.report("IGM: constraints reduced, gradient operators assembled")
        self.ir = IterativeRounding(self.fixedIdcs, self.fixedVals, self.singularIdcs, self.integerIdcs,
                                    self.length, Cfull, self.G2full, N, n, cut.nF, self.seamless,
                                    self.roundSeams, self.rawF2Vec, literal=literal)
        if self.localInjectivity:
            from .injective import locally_injective, run_local_injectivity
            f2n, cidx = self.fullx2Nfunc, np.asarray(self.cut.idx)
            self.success = run_local_injectivity(self.ir, lambda X: locally_injective(f2n, cidx, N, X))
        else:
            self.success = self.ir.run()
        self.fullx = self.ir.fullx
        NF = self.fullx2Nfunc @ self.fullx
        self.nfn = NF.reshape(cut.nV, N)
        self.cfn = np.zeros((raw.nF, 3 * N))
        for j in range(3):
            self.cfn[:, N * j:N * j + N] = self.nfn[cut.idx[:, j]]
        return self.success


def reduce_to_linearly_independent(C):
    """``reduce_to_linearly_independent`` (matrix_ops.h:12-28): the rows SparseQR<COLAMD> on
    C^T keeps, in its order (``oracle.colamd``); explicit zeros stay in the pattern."""
    if C.shape[0] == 0:
        return C
    return _reduce_colamd(C)[1]


class IterativeRounding:
    """``iterative_rounding`` (iter_rounding.h:33-113) with NaiveIntegration (iter_rounding_init.h)
    and SeamlessIntegration (iter_rounding_loop.h), LM from levenberg_marquardt.h."""

    def __init__(self, fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n, nF,
                 seamless, roundSeams, F2, literal=False):
        self.fixedIdcs = np.asarray(fixedIdcs)
        self.fixedVals = np.asarray(fixedVals, dtype=np.float64)
        self.singularIdcs = [int(i) for i in singularIdcs]
        self.integerIdcs = [int(i) for i in integerIdcs]
        self.length, self.G2, self.N, self.n, self.nF = length, G2.tocsr(), N, n, nF
        self.seamless, self.roundSeams, self.F2 = seamless, roundSeams, F2
        self.literal = literal
        self.Cfull = Cfull
        self._naive_setup()

    # -- iter_rounding_init.h:74-232 ---------------------------------------------------
    def _naive_setup(self):
        C = self.Cfull.tocsr()   # explicit zeros stay: their columns take part (iter_rounding_init.h:98-110)
        ncols = C.shape[1]
        uniqueJ = np.unique(C.indices) if C.nnz else np.zeros(0, dtype=np.int64)
        nonPart = np.setdiff1d(np.arange(ncols), uniqueJ)
        if C.shape[0] != 0:
            CSmall = C[:, uniqueJ].toarray()
            USmall, self.crank = fullpiv_lu_kernel(CSmall)
        else:
            USmall = np.zeros((0, 0))
            nonPart = np.arange(ncols)
            uniqueJ = np.zeros(0, dtype=np.int64)
            self.crank = 0
        npn = len(nonPart)
        # UFull = permMat * blockdiag(I, USmall)
        rows = list(nonPart)
        cols = list(range(npn))
        vals = [1.0] * npn
        if USmall.size:
            ui, uj = np.nonzero(USmall)
            rows += list(uniqueJ[ui])
            cols += list(npn + uj)
            vals += list(USmall[ui, uj])
        k = npn + (USmall.shape[1] if USmall.size else 0)
        self.UFull = sp.csr_matrix((vals, (rows, cols)), shape=(ncols, k))
        self.k = k
        self.USmall, self.uniqueJ, self.nonPart = USmall, uniqueJ, nonPart
        X2F = _prune_zeros(self.G2 @ self.UFull)
        self.X2F_naive = X2F
        E = (X2F.T @ X2F * self.length).tocsr()
        f = X2F.T @ self.F2
        constMat = self.UFull[self.fixedIdcs, :]
        KKT = sp.bmat([[E, constMat.T], [constMat, None]], format="csc")
        progress.report(f"IGM: kernel basis (rank {self.crank}) and KKT assembled")
        rhs = np.concatenate([f, self.fixedVals])
        self.E, self.f_rhs, self.constMat = E, f, constMat
        sol = spla.splu(KKT).solve(rhs)
        progress.report(f"IGM: KKT solved ({KKT.shape[0]} rows)")
        self.XSmall = sol[:k]
        self.kkt_residual = np.linalg.norm(KKT @ sol - rhs) / max(np.linalg.norm(rhs), 1e-300)
        self.x0 = self.UFull @ self.XSmall
        self.B = (self.G2 @ self.UFull * self.length).tocsr()  # length * G2 * UFull

    # -- LM (levenberg_marquardt.h:140-231) ---------------------------------------------
    @staticmethod
    def _lm(x0, objective, direction, maxIter, funcTol, fooTol, stats):
        """LMSolver::solve with lambda = 0 and verbose = false.

        ``objective(x)`` -> (E, J or None); ``direction(x)`` -> d solving J^T J d = -J^T E.
        Returns (x, return_value)."""
        prevx = x0.copy()
        E, J = objective(prevx, True)
        it = 0
        x = prevx
        while True:
            rhs = -(J.T @ E)
            foo = np.max(np.abs(rhs)) if rhs.size else 0.0
            if foo < fooTol:
                x = prevx  # the reference only breaks here when verbose
            d = direction(prevx)
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
            it += 1
            prevx = x
            if not it <= maxIter:
                return x, False

    def run(self):
        """iter_rounding.h:33-113."""
        self.stats = {"solves": 0, "rounds": 0, "lm_rounds": 0}
        if not self.seamless:
            self.fullx = self.x0
            return True
        wI, wK, wC = 10e3, 10e3, 1.0  # iter_rounding_init.h:184
        B, F2, G2, UF = self.B, self.F2, self.G2, self.UFull
        k, l1 = self.k, len(F2)
        fix = self.fixedIdcs
        Ufix = UF[fix, :]
        fv = self.fixedVals
        # J = [wI(-B | I); wC(0 | I); wK(Ufix | 0)]
        Jn = sp.bmat([[-wI * B, wI * sp.identity(l1)],
                      [None, wC * sp.identity(l1)],
                      [wK * Ufix, None]], format="csr")

        def obj_n(xf, needJ):
            xs, f = xf[:k], xf[k:]
            E = np.concatenate([wI * (f - B @ xs), wC * (f - F2), wK * (Ufix @ xs - fv)])
            return E, (Jn if needJ else None)

        def dir_n(xf):
            if self.literal:
                E, _ = obj_n(xf, False)
                return spla.splu((Jn.T @ Jn).tocsc()).solve(-(Jn.T @ E))
            xs = self.XSmall
            fs = (wI ** 2 * (B @ xs) + wC ** 2 * F2) / (wI ** 2 + wC ** 2)
            return np.concatenate([xs, fs]) - xf

        XF = np.concatenate([self.XSmall, F2])
        xnaive, _ = self._lm(XF, obj_n, dir_n, 100, 10e-10, 10e-10, self.stats)
        progress.report("IGM: naive LM pass done")
        self.x_naive = xnaive
        self.fullx = UF @ xnaive[:k]
        X_ = xnaive[:k].copy()
        F2_ = xnaive[k:].copy()
        self.F2_ = F2_

        # SeamlessIntegration (iter_rounding_loop.h:102-124)
        wI, wK, wC = 1.0, 10e4, 0.01
        X2F = _prune_zeros(G2 @ UF * self.length)
        A0 = (wI ** 2 * (X2F.T @ X2F) + wC ** 2 * sp.identity(k)).tocsc()
        A0lu = spla.splu(A0)
        progress.report(f"IGM: A0 factorised ({A0.shape[0]} rows)")
        b_int = wI ** 2 * (X2F.T @ F2_)
        leftIdcs = list(self.integerIdcs if self.roundSeams else self.singularIdcs)
        roundedSingulars = False
        fixedI, fixedV = [], []
        Zcols = []
        xCurr = X_.copy()
        xInit = X_.copy()
        success, rounded = True, False
        lm_x = None
        self.decisions = []
        while leftIdcs:
            # initFixedIndices (loop.h:29-74)
            xPrev = xCurr.copy()
            X = UF @ xCurr
            minDif, minIdx = 3276700.0, -1
            for i, li in enumerate(leftIdcs):
                v = X[li]
                dd = abs(v - _cround(v))
                if dd < minDif:
                    minIdx, minDif = i, dd
            pick = leftIdcs[minIdx]
            fixedI.append(pick)
            fixedV.append(int(_cround(X[pick])))
            self.decisions.append((pick, int(_cround(X[pick])), float(minDif)))
            del leftIdcs[minIdx]
            u = UF[pick, :].toarray().ravel()
            Zcols.append(A0lu.solve(u))
            if not leftIdcs and not self.roundSeams and not roundedSingulars:
                leftIdcs = list(self.integerIdcs)
                roundedSingulars = True
            self.stats["rounds"] += 1
            progress.report(f"IGM: rounding round {self.stats['rounds']}")
            if not (minDif > 10e-7):
                continue
            rounded = True
            self.stats["lm_rounds"] += 1
            Uf = UF[fixedI, :].tocsr()
            r = np.array(fixedV, dtype=np.float64)
            Js = sp.bmat([[wI * X2F], [wC * sp.identity(k)], [wK * Uf]], format="csr")

            def obj_s(xs, needJ, xPrev=xPrev, Uf=Uf, r=r, Js=Js):
                E = np.concatenate([wI * (X2F @ xs - F2_), wC * (xs - xPrev), wK * (Uf @ xs - r)])
                return E, (Js if needJ else None)

            Z = np.stack(Zcols, axis=1)
            Msmall = np.eye(len(fixedI)) / wK ** 2 + Uf @ Z
            xt = A0lu.solve(b_int + wC ** 2 * xPrev)
            xstar = xt + Z @ np.linalg.solve(Msmall, r - Uf @ xt)
            # the last LM round's system, for the sizing hook (oracle/sizing.py)
            self.last_round = {"xPrev": xPrev.copy(), "Uf": Uf, "r": r.copy(), "Z": Z, "M": Msmall, "A0lu": A0lu,
                               "X2F": X2F, "wI": wI, "wK": wK, "wC": wC}

            def dir_s(xs, Js=Js, obj_s=obj_s, xstar=xstar):
                if self.literal:
                    E, _ = obj_s(xs, False)
                    return spla.splu((Js.T @ Js).tocsc()).solve(-(Js.T @ E))
                return xstar - xs

            lm_x, _ = self._lm(xInit, obj_s, dir_s, 100, 1e-7, 1e-7, self.stats)
            # post_checking (loop.h:89-100)
            fx = UF @ lm_x
            mx = max(abs(fx[i] - _cround(fx[i])) for i in fixedI)
            xCurr = lm_x.copy()
            xInit = lm_x.copy()
            if not (mx <= 10e-7):
                success = False
                break
        self.fixedI, self.fixedV = fixedI, fixedV
        self.x_final = lm_x
        self.fullx = UF @ lm_x if rounded else self.x0
        return success
