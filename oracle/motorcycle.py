"""Oracle: the motorcycle graph of QGP (CPU, FP64).

Test infrastructure only (see ``oracle/__init__.py``).  Restates the reference's
``core/tmesh/motorcycle.h`` as example_qgp.cpp drives it (``MotorcycleGraph(mesh, uv,
matching, singular)``, example_qgp.cpp:110):

* ``gen_ports`` (motorcycle.h:165-212): per singular vertex (ascending id), per outgoing
  halfedge (Vert::adjHalfs order), the grid directions pointing into the corner or along
  one of its two edges, starting after the first non-inward rotation, sorted by their dot
  product with the corner's first edge (descending), numbered in that order and linked
  cyclically per vertex;
* ``add_segment_init`` / ``add_segment_next`` / ``add_segment`` / ``split_segment``
  (motorcycle.h:104-119, 214-296): every curve grows face by face (all curves one face per
  sweep, in port order) until it strictly crosses a segment of another curve in the same
  face (nearest crossing first); that segment is split there (HitL / HitR on its two
  halves, by the side the crash comes from) and the curve ends (HitB).  Crossing an edge
  turns the direction by polar(1, pi/2 m) with the edge's matching m;
* the helpers ``isccw`` / ``points_into`` / ``orientation`` of common/utilities.h and
  ``get_oppsite_half`` of hmesh/utilities.h.

``cfn`` is the corner uv (complex, corner 3 f + j at vertex F[f, j]); closed meshes.
"""
import cmath
import math

import numpy as np

NONE, HITL, HITR, HITB = 0, 1, 2, 3


def cross(a, b):
    return a.real * b.imag - a.imag * b.real


def dot(a, b):
    return a.real * b.real + a.imag * b.imag


def isccw(a, b, c):
    return cross(b - a, c - a) > 0


def points_into(d, a, b, c):
    return isccw(a, b, a + d) and isccw(a, a + d, c)


def quater_rot(i):
    return (1 + 0j, 1j, -1 + 0j, -1j)[(i % 4 + 4) % 4]


def lerp(a, b, t):
    return a * (1.0 - t) + b * t


def _inv2(m00, m01, m10, m11):
    inv = 1.0 / (m00 * m11 - m10 * m01)
    return m11 * inv, -m01 * inv, -m10 * inv, m00 * inv


def extended_intersection(a, b, c, d):
    m00, m01 = d.real - c.real, -(b - a).real
    m10, m11 = d.imag - c.imag, -(b - a).imag
    if m00 * m11 - m10 * m01 == 0:
        return False, 0.0, 0.0
    i00, i01, i10, i11 = _inv2(m00, m01, m10, m11)
    x0, x1 = -c.real + a.real, -c.imag + a.imag
    return True, i10 * x0 + i11 * x1, i00 * x0 + i01 * x1


def strict_intersection(a, b, c, d):
    ok, rab, rcd = extended_intersection(a, b, c, d)
    return ok and 0 < rab < 1 and 0 < rcd < 1, rab, rcd


class Seg:
    __slots__ = ("curv", "face", "fr", "to", "fr_type", "to_type", "fr_crash", "to_crash")

    def __init__(self, curv, face, fr, to, fr_type=NONE, to_type=NONE, fr_crash=-1, to_crash=-1):
        self.curv, self.face, self.fr, self.to = curv, face, fr, to
        self.fr_type, self.to_type, self.fr_crash, self.to_crash = fr_type, to_type, fr_crash, to_crash

    def same(self, o):
        return self.face == o.face and self.fr == o.fr and self.to == o.to


class MotorcycleGraph:
    def __init__(self, m, cfn, matching, singular, max_sweeps=None):
        if np.any(m.face < 0):
            raise ValueError("motorcycle graph: closed meshes only")
        self.m = m
        self.cfn = [complex(c) for c in cfn]
        self.matching = np.asarray(matching)
        self.ports = self._gen_ports(np.asarray(singular))
        n = len(self.ports)
        self.segs = [[] for _ in range(n)]
        self.cache = [None] * n        # (half, ratio, dir)
        self.done = [False] * n
        for i in range(n):
            self._init(i)
        sweeps = 0
        max_sweeps = max_sweeps or 4 * m.nF + 16
        while not all(self.done):
            for i in range(n):
                self._next(i)
            sweeps += 1
            if sweeps > max_sweeps:
                raise RuntimeError("motorcycle graph: a curve never crashed")

    def _gen_ports(self, singular):
        m, cfn = self.m, self.cfn
        out = []
        counter = 0
        for v in range(m.nV):
            if singular[v] == 0:
                continue
            ports = []
            for h in m.vert_halfs_ccw(v):
                a, b, c = cfn[m.crnr[m.next[h]]], cfn[m.crnr[m.prev[h]]], cfn[m.crnr[h]]
                if cross(b - a, c - a) < 0:
                    raise ValueError("Orientation should be ccw order")
                r = 0
                while r < 4:
                    if not points_into(quater_rot(r), a, b, c) and r > 0:
                        break
                    r += 1
                temps = []
                for i in range(4):
                    d = quater_rot(r - i)
                    if points_into(d, a, b, c) or cmath.phase(d) == cmath.phase(b - a) or \
                            cmath.phase(d) == cmath.phase(c - a):
                        temps.append({"vert": v, "face": int(m.face[h]), "uv": a, "dir": d})
                e = b - a
                temps.sort(key=lambda p: -dot(p["dir"], e))   # stable, descending
                for p in temps:
                    p["id"] = counter
                    counter += 1
                ports += temps
            s = len(ports)
            for i, p in enumerate(ports):
                p["prev"] = ports[(i - 1 + s) % s]["id"]
                p["next"] = ports[(i + 1 + s) % s]["id"]
            out += ports
        return out

    def _init(self, i):
        m = self.m
        p = self.ports[i]
        h = -1
        for hh in m.face_halfs(p["face"]):
            if m.head[hh] != p["vert"] and m.tail[hh] != p["vert"]:
                h = hh
        self._add(i, p["dir"], p["uv"], h)

    def _next(self, i):
        if self.done[i]:
            return
        m, cfn = self.m, self.cfn
        h0, ratio, d0 = self.cache[i]
        h = m.twin[h0]
        k = (1 if m.edge2half[m.edge[h]] == h else -1) * int(self.matching[m.edge[h]])
        uv0 = lerp(cfn[m.crnr[m.next[h]]], cfn[m.crnr[m.prev[h]]], ratio)
        d = cmath.rect(1.0, math.pi / 2 * k) * d0
        self._add(i, d, uv0, self._opposite_half(uv0, d, h))

    def _opposite_half(self, origin, d, fr):
        m, cfn = self.m, self.cfn
        for h in m.face_halfs(m.face[fr]):
            if h == fr:
                continue
            a = cfn[m.crnr[m.prev[h]]] - origin
            b = cfn[m.crnr[m.next[h]]] - origin
            if cross(a, d) * cross(b, d) < 0:
                return h
        raise RuntimeError("motorcycle graph: no opposite half")

    def _add(self, i, d, uv0, h):
        m, cfn = self.m, self.cfn
        uv1, uv2 = cfn[m.crnr[m.prev[h]]], cfn[m.crnr[m.next[h]]]
        ok, r_ab, r_cd = extended_intersection(uv0, uv0 + d, uv1, uv2)
        if not ok:
            raise RuntimeError("motorcycle graph: parallel to the exit edge")
        uv3 = lerp(uv1, uv2, r_cd)
        f = int(m.face[h])
        best = None
        for j, segs in enumerate(self.segs):
            if j == i:
                continue
            for sg in segs:
                if sg.face != f:
                    continue
                hit, ab, cd = strict_intersection(uv0, uv3, sg.fr, sg.to)
                if hit and (best is None or ab < best[0]):
                    best = (ab, cd, sg)
        if best is None:
            self.segs[i].append(Seg(i, f, uv0, uv3))
            self.cache[i] = (h, r_cd, d)
            return
        ab, cd, sg = best
        crash_at = lerp(uv0, uv3, ab)
        self.segs[i].append(Seg(i, f, uv0, crash_at, to_type=HITB, to_crash=sg.curv))
        self.cache[i] = (h, r_cd, d)
        self.done[i] = True
        self._split(sg.curv, sg, self.segs[i][-1], cd)

    def _split(self, j, target, crash, ratio):
        segs = self.segs[j]
        for k, s in enumerate(segs):
            if s.same(target):
                ccw = cross(crash.to - crash.fr, s.to - s.fr) > 0
                uv = lerp(s.fr, s.to, ratio)
                tail = Seg(j, s.face, uv, s.to, fr_type=HITL if ccw else HITR, to_type=s.to_type,
                           fr_crash=crash.curv, to_crash=s.to_crash)
                s.to, s.to_type, s.to_crash = uv, (HITR if ccw else HITL), crash.curv
                segs.insert(k + 1, tail)
                return
