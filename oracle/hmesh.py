"""Oracle: halfedge mesh and per-element geometry (FP64, CPU).

Restates ``hmesh/hmesh.h`` (struct ``Hmesh``, ``dcel``) and libigl v2.5.0's
``igl::edge_topology`` that it calls (``hmesh/hmesh.h:317``).  Test
infrastructure only; see ``oracle/__init__.py``.
"""
import numpy as np

PI = np.pi
TWO_PI = 2.0 * np.pi


def edge_topology(F):
    """libigl v2.5.0 ``igl::edge_topology(V, F, EV, FE, EF)`` (published algorithm).

    Rows (min v, max v, f, i) for every face side i (F(f,i) -> F(f,(i+1)%3))
    are sorted lexicographically; consecutive equal (v1, v2) rows form one
    edge, numbered in sorted order.  EF(e,0) is then made the face in which
    the edge runs EV(e,0) -> EV(e,1) ("the face on the left"), swapping with
    EF(e,1) otherwise (a border edge may end up with EF(e,0) = -1).
    Called at ``hmesh/hmesh.h:317``.
    """
    F = np.asarray(F, dtype=np.int64)
    nF = F.shape[0]
    a = F.reshape(-1)
    b = F[:, [1, 2, 0]].reshape(-1)
    v1 = np.minimum(a, b)
    v2 = np.maximum(a, b)
    fi = np.repeat(np.arange(nF), 3)
    si = np.tile(np.arange(3), nF)
    order = np.lexsort((si, fi, v2, v1))
    s1, s2, sf, ss = v1[order], v2[order], fi[order], si[order]
    n = len(order)
    same_next = np.zeros(n, dtype=bool)
    same_next[:-1] = (s1[:-1] == s1[1:]) & (s2[:-1] == s2[1:])
    # greedy pairing exactly as the reference loop: row i pairs with i+1
    # when their keys agree, and the pair is skipped as a unit.
    EV, EF, FE = [], [], -np.ones((nF, 3), dtype=np.int64)
    i = 0
    while i < n:
        if i < n - 1 and same_next[i]:
            e = len(EV)
            EV.append((s1[i], s2[i]))
            EF.append([sf[i], sf[i + 1]])
            FE[sf[i], ss[i]] = e
            FE[sf[i + 1], ss[i + 1]] = e
            if i + 2 < n and same_next[i + 1]:
                raise ValueError("edge_topology: mesh is not edge-manifold")
            i += 2
        else:
            e = len(EV)
            EV.append((s1[i], s2[i]))
            EF.append([sf[i], -1])
            FE[sf[i], ss[i]] = e
            i += 1
    EV = np.array(EV, dtype=np.int64).reshape(-1, 2)
    EF = np.array(EF, dtype=np.int64).reshape(-1, 2)
    # orientation: first face contains EV0 -> EV1
    for e in range(EV.shape[0]):
        fid = EF[e, 0]
        flip = True
        for j in range(3):
            if F[fid, j] == EV[e, 0] and F[fid, (j + 1) % 3] == EV[e, 1]:
                flip = False
        if flip:
            EF[e, 0], EF[e, 1] = EF[e, 1], EF[e, 0]
    return EV, FE, EF


def _normalize_rows(x):
    n = np.sqrt((x * x).sum(axis=1))
    return x / n[:, None]


class Hmesh:
    """Restatement of ``struct Hmesh`` and its constructor (``hmesh/hmesh.h:96-147,313-569``).

    Halfedge numbering follows ``dcel`` (``hmesh.h:240-311``): for each edge
    in edge_topology order, the halfedge of EF(e,0) then of EF(e,1) get the
    next ids (so 2e/2e+1 on closed meshes); boundary halfedges are appended
    loop by loop (``hmesh.h:391-416``).
    """

    def __init__(self, V, F):
        V = np.asarray(V, dtype=np.float64)
        F = np.asarray(F, dtype=np.int64)
        self.pos = V
        self.idx = F
        EV, FE, EF = edge_topology(F)
        self.edge2vert, self.face2edge, self.edge2face = EV, FE, EF
        nV, nF, nE = V.shape[0], F.shape[0], EV.shape[0]
        self.nV, self.nF, self.nE, self.nH, self.nC = nV, nF, nE, 2 * nE, 3 * nF
        self.nEularChars = nV - nE + nF
        nH = 2 * nE
        self._topology(EV, FE, EF)
        self._geometry()

    # -- hmesh.h:240-311 (dcel) and hmesh.h:364-442 ---------------------------------
    def _topology(self, EV, FE, EF):
        F = self.idx
        nV, nF, nE, nH = self.nV, self.nF, self.nE, self.nH
        EFi = -np.ones((nE, 2), dtype=np.int64)
        for k in range(2):
            m = EF[:, k] != -1
            e = np.nonzero(m)[0]
            f = EF[e, k]
            for j in range(3):
                hit = FE[f, j] == e
                EFi[e[hit], k] = j
        # dcel numbering (hmesh.h:256-262)
        EH = -np.ones((nE, 2), dtype=np.int64)
        numH = 0
        for i in range(nE):
            if EF[i, 0] != -1:
                EH[i, 0] = numH
                numH += 1
            if EF[i, 1] != -1:
                EH[i, 1] = numH
                numH += 1
        HE = np.zeros(numH, dtype=np.int64)
        HV = np.zeros(numH, dtype=np.int64)
        VH = np.zeros(nV, dtype=np.int64)
        HF = np.zeros(numH, dtype=np.int64)
        FH = np.zeros((nF, 3), dtype=np.int64)
        twinH = -np.ones(numH, dtype=np.int64)
        for i in range(nE):
            for k in range(2):
                if EH[i, k] != -1:
                    HE[EH[i, k]] = i
            # vertex -> half: the last assignment in edge order wins (hmesh.h:272-281)
            if EH[i, 0] != -1:
                HV[EH[i, 0]] = EV[i, 0]
                VH[EV[i, 0]] = EH[i, 0]
            if EH[i, 1] != -1:
                HV[EH[i, 1]] = EV[i, 1]
                VH[EV[i, 1]] = EH[i, 1]
            if EH[i, 0] != -1 and EH[i, 1] != -1:
                twinH[EH[i, 0]] = EH[i, 1]
                twinH[EH[i, 1]] = EH[i, 0]
            for k in range(2):
                if EF[i, k] != -1:
                    FH[EF[i, k], EFi[i, k]] = EH[i, k]
                    HF[EH[i, k]] = EF[i, k]
        nextH = np.zeros(numH, dtype=np.int64)
        prevH = np.zeros(numH, dtype=np.int64)
        for j in range(3):
            nextH[FH[:, j]] = FH[:, (j + 1) % 3]
            prevH[FH[:, (j + 1) % 3]] = FH[:, j]

        self.vert2half = VH.copy()
        self.edge2half = EH[:, 0].copy()
        self.face2half = FH[:, 0].copy()
        nxt = -np.ones(nH, dtype=np.int64)
        prv = -np.ones(nH, dtype=np.int64)
        twin = -np.ones(nH, dtype=np.int64)
        head = -np.ones(nH, dtype=np.int64)
        tail = -np.ones(nH, dtype=np.int64)
        edge = -np.ones(nH, dtype=np.int64)
        face = -np.ones(nH, dtype=np.int64)
        head[:numH] = HV[nextH]
        tail[:numH] = HV
        edge[:numH] = HE
        face[:numH] = HF
        nxt[:numH] = nextH
        prv[:numH] = prevH
        twin[:numH] = twinH

        # boundary loops (hmesh.h:391-416)
        loop2half = []
        nP = 3
        iB = nF * nP
        for iH in range(nF * nP):
            if twin[iH] != -1:
                continue
            jH = iH
            jB = iB
            while True:
                tail[iB] = head[jH]
                head[iB] = tail[jH]
                edge[iB] = edge[jH]
                prv[iB] = iB - 1
                nxt[iB] = iB + 1
                twin[jH] = iB
                twin[iB] = jH
                jH = prv[jH]
                done = False
                while twin[jH] != -1:
                    if jH == iH:
                        done = True
                        break
                    jH = prv[twin[jH]]
                if done:
                    break
                iB += 1
            prv[jB] = iB
            nxt[iB] = jB
            loop2half.append(jB)
            iB += 1
        self.loop2half = np.array(loop2half, dtype=np.int64)
        self.nL = len(loop2half)
        for iE in range(nE):
            if self.edge2half[iE] == -1:
                self.edge2half[iE] = twin[EH[iE, 1]]
        self.next, self.prev, self.twin = nxt, prv, twin
        self.head, self.tail, self.edge, self.face = head, tail, edge, face

        # corners (hmesh.h:430-442): corner (f, it) -> the halfedge of f not touching F(f,it)
        self.crnr2half = -np.ones(3 * nF, dtype=np.int64)
        self.crnr = -np.ones(nH, dtype=np.int64)
        h0 = self.face2half
        hs = [h0, nxt[h0], prv[h0]]
        for it in range(3):
            vid = F[:, it]
            for h in hs:
                m = (tail[h] != vid) & (head[h] != vid)
                cid = np.arange(nF)[m] * 3 + it
                self.crnr2half[cid] = h[m]
                self.crnr[h[m]] = cid

        # boundary vertices (hmesh.h:549-555)
        hb = self.edge2half
        eb = (face[hb] == -1) | (face[twin[hb]] == -1)
        self.isBV = np.zeros(nV, dtype=bool)
        self.isBV[tail[hb[eb]]] = True
        self.isBV[head[hb[eb]]] = True
        self.edge_is_boundary = eb

    # -- hmesh.h:461-568 ---------------------------------------------------------------
    def _geometry(self):
        P = self.pos
        nF, nV, nH = self.nF, self.nV, self.nH
        head, tail, nxt, prv, twin, face = self.head, self.tail, self.next, self.prev, self.twin, self.face
        h0 = self.face2half
        x = P[head[h0]] - P[tail[h0]]
        t = -(P[head[prv[h0]]] - P[tail[prv[h0]]])
        n = np.cross(x, t)
        self.faceBasisX = _normalize_rows(x)
        self.faceBasisY = -_normalize_rows(np.cross(x, n))
        self.faceNormal = _normalize_rows(n)
        self.faceArea = np.sqrt((n * n).sum(axis=1)) * 0.5
        self.baryCenter = (P[tail[h0]] + P[tail[nxt[h0]]] + P[tail[nxt[nxt[h0]]]]) / 3.0

        # heArgOnFace (hmesh.h:481-489): exterior angles accumulated from face2half,
        # skipping the halfedge whose direction is the face basis x
        self.heArgOnFace = np.zeros(nH)
        s = np.zeros(nF)
        h = h0.copy()
        for _ in range(3):
            v1 = _normalize_rows(P[head[h]] - P[tail[h]])
            v2 = _normalize_rows(P[head[prv[h]]] - P[tail[prv[h]]])
            d = np.sqrt(((v1 - self.faceBasisX) ** 2).sum(axis=1))
            add = np.arccos((v1 * v2).sum(axis=1))
            s = s + np.where(d > 1e-10, add, 0.0)
            self.heArgOnFace[h] = s
            h = nxt[h]

        # dihedral angle per halfedge (hmesh.h:508-519): atan2(v . (n1 x n2), n1 . n2), v the
        # unit halfedge direction, n1 / n2 the normals of its face and of its twin's face; 0 on
        # boundary edges
        self.dihedralArg = np.zeros(nH)
        hh = np.nonzero((face >= 0) & (face[twin] >= 0))[0]
        n1 = self.faceNormal[face[hh]]
        n2 = self.faceNormal[face[twin[hh]]]
        vh = P[head[hh]] - P[tail[hh]]
        vh = vh / np.sqrt((vh * vh).sum(axis=1))[:, None]
        self.dihedralArg[hh] = np.arctan2((vh * np.cross(n1, n2)).sum(axis=1), (n1 * n2).sum(axis=1))

        # vertex normals / bases (hmesh.h:492-506)
        vn = np.zeros((nV, 3))
        for j in range(3):
            np.add.at(vn, self.idx[:, j], self.faceNormal * self.faceArea[:, None])
        self.vertNormal = _normalize_rows(vn)
        vh = self.vert2half
        vv = P[head[vh]] - P
        vv = vv - (vv * self.vertNormal).sum(axis=1)[:, None] * self.vertNormal
        self.vertBasisX = _normalize_rows(vv)
        self.vertBasisY = np.cross(self.vertNormal, self.vertBasisX)

        # cotangents (hmesh.h:522-533)
        interior = face >= 0
        hc = np.zeros(nH)
        hi = np.nonzero(interior)[0]
        vn_ = P[head[nxt[hi]]] - P[tail[nxt[hi]]]
        vp_ = -(P[head[prv[hi]]] - P[tail[prv[hi]]])
        hc[hi] = (vp_ * vn_).sum(axis=1) / np.sqrt((np.cross(vp_, vn_) ** 2).sum(axis=1))
        self.halfCotan = hc
        self.edgeCotan = (hc[self.edge2half] + hc[twin[self.edge2half]]) * 0.5

        # angle defect (hmesh.h:558-568): sum of angles between consecutive
        # outgoing halfedges h and prev(h).twin() around each vertex
        vec = P[head] - P[tail]
        v1 = _normalize_rows(vec)
        v2 = _normalize_rows(vec[twin[prv]])
        with np.errstate(invalid="ignore"):  # NaN on cut-mesh border halfedges, as in the reference
            ang = np.arccos((v1 * v2).sum(axis=1))
        asum = np.zeros(nV)
        np.add.at(asum, tail, ang)
        self.angleSum = asum
        self.angleDefect = np.where(self.isBV, PI - asum, TWO_PI - asum)

    # -- helpers mirroring the reference's element accessors ---------------------------
    def is_canonical(self, h):
        """``Half::isCanonical`` (hmesh.h:207)."""
        return self.edge2half[self.edge[h]] == h

    def vert_halfs_ccw(self, v, start=None):
        """``Vert::adjHalfs(ccw=true)`` (hmesh.h:160): h -> h.prev().twin()."""
        h = self.vert2half[v] if start is None else start
        out = [h]
        while True:
            h = self.twin[self.prev[h]]
            if h == out[0]:
                return out
            out.append(h)

    def vert_halfs_cw(self, v, start=None):
        """``Vert::adjHalfs(ccw=false)`` (hmesh.h:160): h -> h.twin().next()."""
        h = self.vert2half[v] if start is None else start
        out = [h]
        while True:
            h = self.next[self.twin[h]]
            if h == out[0]:
                return out
            out.append(h)

    def face_halfs(self, f):
        """``Face::adjHalfs()`` (hmesh.h:165)."""
        h = self.face2half[f]
        return [h, self.next[h], self.next[self.next[h]]]

    def half_vec(self, h):
        return self.pos[self.head[h]] - self.pos[self.tail[h]]
