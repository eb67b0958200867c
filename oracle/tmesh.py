"""Oracle: the T-mesh of QGP built from the motorcycle graph (CPU).

Test infrastructure only (see ``oracle/__init__.py``).  Restates the reference's
``core/tmesh/tmesh.h`` (Tmesh, Tquad, Thalf, Tedge; example_qgp.cpp:111):

* T-edges (tmesh.h:158-169): each curve, in order, is cut after every segment whose end is a
  hit vertex; T-edge s runs from the first segment after the previous cut to that segment,
  with half-edges 2 s (canonical, along the curve) and 2 s + 1 (against it);
* T-quads (tmesh.h:95-118, 171-178): from the first unvisited half-edge, follow
  ``choose_next_thalf`` (tmesh.h:262-305) until the start comes back, counting a side change
  at every turn, then rotate so that the list does not start in the middle of a side;
* ``th2quad`` / ``th2side`` / ``th2iter`` and ``th2sing`` (the half-edge leaving each
  singular vertex, tmesh.h:185-209).

A motorcycle graph is ``oracle.motorcycle.MotorcycleGraph``; a segment is identified by
(curve, index) and compared as the reference's Msgmt::operator== does (face and both end
points, exactly).
"""
from .motorcycle import HITB, HITR, NONE, cross

_EQ_TOL = 1e-10


class Tmesh:
    def __init__(self, g):
        self.g = g
        segs = g.segs
        self.tedges = []       # (curve, first segment, last segment)
        for c, ss in enumerate(segs):
            start = 0
            for k, s in enumerate(ss):
                if s.to_type != NONE:
                    self.tedges.append((c, start, k))
                    if k + 1 < len(ss):
                        start = k + 1
        nth = 2 * len(self.tedges)
        self.tquads = []       # (thids, sides)
        visit = [False] * nth
        while not all(visit):
            b = visit.index(False)
            q = self._quad(b)
            self.tquads.append(q)
            for t in q[0]:
                visit[t] = True
        self.th2quad = [0] * nth
        self.th2side = [0] * nth
        self.th2iter = [0] * nth
        for qi, (thids, sides) in enumerate(self.tquads):
            for j, t in enumerate(thids):
                self.th2quad[t], self.th2side[t], self.th2iter[t] = qi, sides[j], j
        self.th2sing = [-1] * nth
        for c, ss in enumerate(segs):
            first = ss[0]
            t = next(i for i in range(nth) if self._seg(self._seg_fr(i)).same(first))
            self.th2sing[t] = g.ports[c]["vert"]

    # half-edge helpers: canonical iff even
    def _edge(self, t):
        return self.tedges[t // 2]

    def _seg(self, ref):
        c, k = ref
        return self.g.segs[c][k]

    def _seg_fr(self, t):
        c, a, b = self._edge(t)
        return (c, a)       # the T-edge's first segment (Tedge::seg_fr)

    def _find_th(self, seg, canonical):
        for t in range(2 * len(self.tedges)):
            if (t % 2 == 0) != canonical:
                continue
            c, a, b = self._edge(t)
            if self.g.segs[c][a].same(seg) or self.g.segs[c][b].same(seg):
                return t
        raise RuntimeError("thalf not found")

    def _choose_next(self, t):
        g = self.g
        c, a, b = self._edge(t)
        sfr, sto = g.segs[c][a], g.segs[c][b]
        if t % 2 == 0:
            if sto.to_type == HITB:
                uv2, dif = sto.to, sto.to - sto.fr
                for ms in g.segs[sto.to_crash]:
                    if ms.face != sto.face:
                        continue
                    if abs(ms.fr - uv2) < _EQ_TOL and cross(dif, ms.to - ms.fr) > 0:
                        return self._find_th(ms, True), True
                    if abs(ms.to - uv2) < _EQ_TOL and cross(dif, -(ms.to - ms.fr)) > 0:
                        return self._find_th(ms, False), True
                raise RuntimeError("thalf not found")
            if sto.to_type == HITR:
                return self._find_th(g.segs[sto.to_crash][-1], False), True
            return t + 2, False
        if a == 0:
            prev_curve = g.ports[c]["prev"]
            return self._find_th(g.segs[prev_curve][0], True), True
        if sfr.fr_type == HITR:
            return self._find_th(g.segs[sfr.fr_crash][-1], False), True
        return t - 2, False

    def _quad(self, b):
        thids, sides = [], []
        side, cur = 0, b
        while True:
            thids.append(cur)
            sides.append(side)
            nxt, turn = self._choose_next(cur)
            if turn:
                side = (side + 1) % 4
            cur = nxt
            if cur == b:
                break
            if len(thids) > 4 * len(self.tedges) + 8:
                raise RuntimeError("T-quad does not close")
        if sides[0] == sides[-1]:
            i = sides[0]
            n = 0
            while n < len(sides) and sides[n] == i:
                n += 1
            sides = sides[n:] + sides[:n]
            thids = thids[n:] + thids[:n]
        if sides[0] == sides[-1]:
            raise RuntimeError("T-quad with a single side")
        return thids, sides
