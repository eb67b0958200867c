"""Oracle: QGP quantization of the T-mesh edge lengths (CPU).

Test infrastructure only (see ``oracle/__init__.py``).  Restates the reference's
``core/quantization/*.h`` on an ``oracle.tmesh.Tmesh``:

* ``compute_weight`` (quantization_weight.h), ``compute_constraint``
  (quantization_constraint.h), ``compute_validation`` (quantization_validation.h);
* ``gen_basis_loop`` (quantization_basisloop.h): a best-first search over a
  ``std::priority_queue``; its pop order among entries the comparator does not separate is
  libstdc++'s heap order, so the queue here is a restatement of libstdc++'s
  ``push_heap`` / ``pop_heap`` (sift-up; hole to a leaf along the larger child, then sift-up);
* ``construct_generating_vectors`` (quantization_evaluation.h) with the reference's row
  choice for ``reduce_to_linearly_independent`` (SparseQR<COLAMD>, ``oracle.colamd``);
* ``compute_quantization`` (quantization.h:28-122), with ties of the weight sort kept in
  (edge, sign) order, and the ideal lengths R of example_qgp.cpp:112-123.
"""
import math

import numpy as np

import scipy.sparse as sp

from .colamd import reduce_to_linearly_independent
from .motorcycle import HITB, HITL, HITR, NONE


def compute_weight(ideal, curr, n_tedges, minus=False):
    if minus and curr == 0:
        return 1e9
    d = (ideal - curr) * (-1.0 if minus else 1.0)
    n = float(n_tedges)
    if d >= 1.0:
        return 1.0 / (d + 1.0)
    if d < 0.0:
        return n * n / (1.0 - d)
    return n / (d + 1.0)


class _Heap:
    """std::priority_queue with comparator less(a, b): top = the last element under less."""

    def __init__(self, less):
        self.c, self.less = [], less

    def _push_up(self, hole, top, value):
        c, less = self.c, self.less
        parent = (hole - 1) // 2
        while hole > top and less(c[parent], value):
            c[hole] = c[parent]
            hole = parent
            parent = (hole - 1) // 2
        c[hole] = value

    def push(self, x):
        self.c.append(x)
        self._push_up(len(self.c) - 1, 0, x)

    def top(self):
        return self.c[0]

    def pop(self):
        c, less = self.c, self.less
        n = len(c)
        if n > 1:
            value = c[n - 1]
            c[n - 1] = c[0]
            ln = n - 1                  # __adjust_heap(first, 0, len - 1, value)
            hole, top, child = 0, 0, 0
            while child < (ln - 1) // 2:
                child = 2 * (child + 1)
                if less(c[child], c[child - 1]):
                    child -= 1
                c[hole] = c[child]
                hole = child
            if (ln & 1) == 0 and child == (ln - 2) // 2:
                child = 2 * (child + 1)
                c[hole] = c[child - 1]
                hole = child - 1
            self._push_up(hole, top, value)
        c.pop()

    def __bool__(self):
        return bool(self.c)


class Quantization:
    def __init__(self, tm):
        self.tm = tm
        g = tm.g
        self.nE = len(tm.tedges)
        # ideal lengths: uv length of each T-edge's segments
        self.R = np.zeros(self.nE)
        for i, (c, a, b) in enumerate(tm.tedges):
            self.R[i] = sum(abs(s.to - s.fr) for s in g.segs[c][a:b + 1])

    # ---- T-mesh navigation (Thalf: 2 e canonical, 2 e + 1 reversed) ----
    def _next(self, t):
        q = self.tm.tquads[self.tm.th2quad[t]][0]
        return q[(self.tm.th2iter[t] + 1) % len(q)]

    def _prev(self, t):
        q = self.tm.tquads[self.tm.th2quad[t]][0]
        return q[(self.tm.th2iter[t] - 1 + len(q)) % len(q)]

    @staticmethod
    def _twin(t):
        return t ^ 1

    def _adj(self, t):
        """Thalf::adj_thalfs (tmesh.h:238-259)."""
        g = self.tm.g
        c, a, b = self.tm.tedges[t // 2]
        # sg_to: the T-edge's last segment for the canonical half, its first for the other
        if t % 2 == 0:
            typ = g.segs[c][b].to_type
        else:
            typ = g.segs[c][a].fr_type
        nx = self._next(t)
        if typ == HITR:
            return [nx, self._next(self._twin(nx))]
        if typ in (HITB, HITL):
            return [nx, self._twin(self._prev(self._twin(t)))]
        return []

    def valid_from(self, X, t0):
        visited = [False] * self.nE
        q = [t0]
        while q:
            t = q.pop(0)
            c, a, b = self.tm.tedges[t // 2]
            if t % 2 == 1 and a == 0:
                return False
            for p in self._adj(t):
                e = p // 2
                if not visited[e] and X[e] == 0:
                    visited[e] = True
                    q.append(p)
        return True

    def valid(self, X):
        if not np.all(X > 0):
            return False
        return all(self.valid_from(X, t) for t in range(2 * self.nE) if self.tm.th2sing[t] > -1)

    def opposite_sides_equal(self, X):
        for thids, sides in self.tm.tquads:
            for i in range(2):
                a = sum(int(X[t // 2]) for t, s in zip(thids, sides) if s == i)
                b = sum(int(X[t // 2]) for t, s in zip(thids, sides) if s == i + 2)
                if a != b:
                    return False
        return True

    def basis_loop(self, bgn, less):
        tm = self.tm
        heap = _Heap(less)
        m = {}
        visited = []
        counter = 0
        heap.push((0, 0.0, bgn))
        while True:
            ln, _, prev = heap.top()
            heap.pop()
            thids, sides = tm.tquads[tm.th2quad[prev]]
            side = tm.th2side[prev]
            pairs = [t for t, s in zip(thids, sides) if s == (side + 2) % 4]
            for p in pairs:
                curr = p ^ 1
                if curr == bgn and counter > 0:
                    res = [prev]
                    idx = prev
                    while True:
                        idx = m.get(idx, 0)
                        res.append(idx)
                        if idx == bgn:
                            return res
                on_path = False
                t = prev
                while t != bgn:
                    if t == curr:
                        on_path = True
                        break
                    t = m.get(t, 0)
                if not on_path and curr not in visited:
                    heap.push((ln - 1, float(self.R[curr // 2]), curr))
                    m.setdefault(curr, prev)
                    visited.append(curr)
            counter += 1
            if not heap:
                raise RuntimeError("gen_basis_loop: no loop back to the start")

    def generators(self, less):
        rows = []
        for t in range(0, 2 * self.nE, 2):
            r = np.zeros(self.nE)
            for u in self.basis_loop(t, less):
                r[u // 2] += 1.0
            rows.append(r)
        G = np.array(rows)
        keep, _ = reduce_to_linearly_independent(sp.csr_matrix(G))   # quantization_evaluation.h:86
        return G[keep].T            # generators as columns

    def quantize(self):
        nE, R = self.nE, self.R
        less1 = lambda a, b: a[0] / max(a[1], 1e-9) < b[0] / max(b[1], 1e-9)
        less2 = lambda a, b: a[0] < b[0]
        G = self.generators(less1)
        X = np.zeros(nE)

        def err(Y):
            s = 0.0
            for i in range(nE):
                d = Y[i] / R[i] - 1.0
                s += d * d
            return math.sqrt(s)

        while np.any(X == 0):
            mn, th = 1e9, 0
            for t in range(2 * nE):
                e = t // 2
                if X[e] != 0:
                    continue
                w = compute_weight(R[e], X[e], nE)
                if w < mn:
                    mn, th = w, t
            cols = [c for c in range(G.shape[1]) if G[th // 2, c] != 0]
            if not cols:
                raise RuntimeError("quantization: a T-edge no generator covers")
            X = X + G[:, cols[0]]
            if self.valid(X):
                break
        e = err(X)
        counter = 0
        while counter < 10:
            prev_e = e
            es = []
            for j in range(nE):
                es.append((compute_weight(R[j], X[j], nE, False), j))
                es.append((compute_weight(R[j], X[j], nE, True), j))
            es.sort(key=lambda x: x[0])
            moved = False
            for _, ei in es:
                for c in range(G.shape[1]):
                    if G[ei, c] == 0:
                        continue
                    x1, x2 = X + G[:, c], X - G[:, c]
                    n1, n2 = err(x1), err(x2)
                    if n1 <= e and self.valid(x1):
                        X, e, moved = x1, n1, True
                        break
                    if n2 <= e and self.valid(x2):
                        X, e, moved = x2, n2, True
                        break
                if moved:
                    break
            counter = counter + 1 if prev_e == e else 0
        G2 = self.generators(less2)
        for t in range(2 * nE):
            for c in range(G2.shape[1]):
                if G2[t // 2, c] == 0:
                    continue
                x1, x2 = X + G2[:, c], X - G2[:, c]
                n1, n2 = err(x1), err(x2)
                if n1 <= e and np.all(x1 >= 0) and self.valid(x1):
                    X, e = x1, n1
                if n2 <= e and np.all(x2 >= 0) and self.valid(x2):
                    X, e = x2, n2
        return X
