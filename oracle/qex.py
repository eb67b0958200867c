"""Oracle: QEx quad extraction from a seamless parameterization (CPU, FP64).

Test infrastructure only (see ``oracle/__init__.py``).  Restates the reference's
``core/qex`` headers as example_qex.cpp drives them (example_qex.cpp:66-140):

* ``compute_trs_matrix`` / ``orientation`` / ``nearby_grid`` (qex/common.h:73-150);
* ``sanitization`` (qex/sanitization.h:9-57);
* ``generate_q_vert`` (qex/gen_q_vert.h:7-56);
* ``generate_vqvert_qport`` / ``generate_eqvert_qport`` / ``generate_fqvert_qport``
  (qex/gen_q_port.h:6-107);
* ``generate_q_edge`` with ``predict_extrinsic_collinear`` / ``pick_next_half``
  (qex/gen_q_edge.h:7-119);
* ``generate_q_faces`` (qex/gen_q_face.h:7-39);
* the helpers of common/utilities.h (``get_quater_rot``, ``lerp``, ``get_minmax_int``,
  ``find_strict_intersection``, ``calc_coefficient``) and hmesh/utilities.h
  (``conversion_2d_3d``), with their 2x2 inverses written out as Eigen's
  ``compute_inverse_size2`` evaluates them (cofactors times 1/det, then the product).

``cfn`` is the corner uv as complex numbers, corner 3 f + j at vertex F[f, j]; closed
meshes only (the reference throws on a boundary crossing).  Where the reference asserts,
this raises.
"""
import math

import numpy as np

EPS = 1e-9


def cround(x):
    """std::round: halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else math.ceil(x - 0.5)


def c_mod(a, n):
    """C++ integer % (truncating)."""
    return int(math.fmod(a, n))


def quater_rot(i):
    return (1 + 0j, 1j, -1 + 0j, -1j)[(i % 4 + 4) % 4]


def orientation(pa, pb, pc):
    acx = pa.real - pc.real
    bcx = pb.real - pc.real
    acy = pa.imag - pc.imag
    bcy = pb.imag - pc.imag
    return acx * bcy - acy * bcx


def is_collinear(pa, pb, pc):
    return abs(orientation(pa, pb, pc)) < EPS


def is_points_into(p1, p2, p3, uv):
    return orientation(p1, p2, uv) > EPS and orientation(p1, p3, uv) < EPS


def is_inside_triangle(pa, pb, pc, uv):
    return orientation(pa, pb, uv) > EPS and orientation(pb, pc, uv) > EPS and orientation(pc, pa, uv) > EPS


def nearby_grid1(a):
    return complex(cround(a.real), cround(a.imag))


def _nearby_x(x, d):
    if d == 0:
        return x
    f = abs(math.fmod(x, 1.0))
    if f < EPS or 1 - f < EPS:
        x += d * EPS
    return float(math.ceil(x)) if d > 0 else float(math.floor(x))


def nearby_grid(uv, d):
    return complex(_nearby_x(uv.real, d.real), _nearby_x(uv.imag, d.imag))


def dot(a, b):
    return a.real * b.real + a.imag * b.imag


def lerp(a, b, t):
    return a * (1.0 - t) + b * t


def _inv2(m00, m01, m10, m11):
    det = m00 * m11 - m10 * m01
    inv = 1.0 / det
    return m11 * inv, -m01 * inv, -m10 * inv, m00 * inv, det


def calc_coefficient(origin, p1, p2, target):
    v1, v2, d = p1 - origin, p2 - origin, target - origin
    i00, i01, i10, i11, _ = _inv2(v1.real, v2.real, v1.imag, v2.imag)
    return complex(i00 * d.real + i01 * d.imag, i10 * d.real + i11 * d.imag)


def find_strict_intersection(a, b, c, d):
    """(ok, ratio_a2b, ratio_c2d) of find_strict_intersection (common/utilities.h)."""
    m00, m01 = d.real - c.real, -(b - a).real
    m10, m11 = d.imag - c.imag, -(b - a).imag
    if m00 * m11 - m10 * m01 == 0:
        return False, 0.0, 0.0
    i00, i01, i10, i11, _ = _inv2(m00, m01, m10, m11)
    x0, x1 = -c.real + a.real, -c.imag + a.imag
    vx, vy = i00 * x0 + i01 * x1, i10 * x0 + i11 * x1
    rab, rcd = vy, vx
    return (rab > 0 and rab < 1 and rcd > 0 and rcd < 1), rab, rcd


def _dot3(a, b):
    return float(a[0]) * float(b[0]) + float(a[1]) * float(b[1]) + float(a[2]) * float(b[2])


def _normalized(v):
    """Eigen normalized(): unchanged when the norm is 0."""
    n2 = _dot3(v, v)
    return v / math.sqrt(n2) if n2 > 0 else v


class Qport:
    __slots__ = ("idx", "vid", "eid", "fid", "uv", "dir", "pos", "connected", "next_id", "prev_id")

    def __init__(self, idx, vid, eid, fid, uv, d, pos):
        self.idx, self.vid, self.eid, self.fid = idx, vid, eid, fid
        self.uv, self.dir, self.pos = uv, d, pos
        self.connected, self.next_id, self.prev_id = False, -1, -1


class Qex:
    """The QEx stages on an oracle Hmesh ``m`` (closed) with corner uv ``cfn``."""

    def __init__(self, m):
        if np.any(m.face < 0):
            raise ValueError("qex: closed meshes only")
        self.m = m

    def crnr_vert(self, c):
        m = self.m
        return m.tail[m.prev[m.crnr2half[c]]]

    def conversion_2d_3d(self, f, cfn, uv):
        m = self.m
        h0 = m.face2half[f]
        c1, c2, c3 = m.crnr[h0], m.crnr[m.next[h0]], m.crnr[m.prev[h0]]
        o3 = m.pos[self.crnr_vert(c1)]
        p1 = m.pos[self.crnr_vert(c2)]
        p2 = m.pos[self.crnr_vert(c3)]
        c = calc_coefficient(cfn[c1], cfn[c2], cfn[c3], uv)
        return o3 + (p1 - o3) * c.real + (p2 - o3) * c.imag

    def trs(self, cfn, matching, N):
        """compute_trs_matrix (qex/common.h:73-94)."""
        m = self.m
        heR = np.zeros(m.nH, complex)
        heT = np.zeros(m.nH, complex)
        for h in range(m.nH):
            uv1 = cfn[m.crnr[m.next[h]]]
            uv2 = cfn[m.crnr[m.prev[m.twin[h]]]]
            k = (1 if m.edge2half[m.edge[h]] == h else -1) * int(matching[m.edge[h]])
            k = c_mod(N + c_mod(k, N), N) if k < 0 else c_mod(k, N)
            r = quater_rot(k)
            t = uv2 - r * uv1
            heR[h] = r
            heT[h] = complex(cround(t.real), cround(t.imag))
        return heR, heT

    def sanitization(self, matching, singular, N, cfn):
        """sanitization (qex/sanitization.h:9-57); returns the sanitized copy of cfn."""
        m = self.m
        cfn = np.array(cfn, dtype=complex)
        heR, heT = self.trs(cfn, matching, N)
        for v in range(m.nV):
            hs = m.vert_halfs_ccw(v)
            mx = 0.0
            for h in hs:
                uv = cfn[m.crnr[m.next[h]]]
                mx = max(mx, abs(uv.real), abs(uv.imag))
            for i, h in enumerate(hs):
                cc = m.crnr[m.next[h]]
                if i == 0:
                    uv = cfn[cc]
                    if singular[v]:
                        cfn[cc] = complex(cround(uv.real), cround(uv.imag))
                    else:
                        delta = 2.0 ** math.log2(mx) if mx > 0 else 0.0
                        sign = complex(int(0 < uv.real) - int(uv.real < 0), int(0 < uv.imag) - int(uv.imag < 0))
                        ds = complex(delta * sign.real, delta * sign.imag)
                        cfn[cc] = (cfn[cc] + ds) - ds
                else:
                    cp = m.crnr[m.prev[m.twin[h]]]
                    cfn[cc] = heR[m.twin[h]] * cfn[cp] + heT[m.twin[h]]
        return cfn

    def q_verts(self, cfn):
        """generate_q_vert (qex/gen_q_vert.h): lists of (uv, pos, sid)."""
        m = self.m
        vq, eq, fq = [], [], []
        for v in range(m.nV):
            uv = cfn[m.crnr[m.next[m.vert2half[v]]]]
            x = math.fmod(abs(uv.real), 1.0)
            y = math.fmod(abs(uv.imag), 1.0)
            if (x < EPS or 1 - x < EPS) and (y < EPS or 1 - y < EPS):
                vq.append((complex(x, y), m.pos[v].copy(), v))
        for e in range(m.nE):
            h = m.edge2half[e]
            p1, p2 = m.pos[m.tail[h]], m.pos[m.head[h]]
            uv1, uv2 = cfn[m.crnr[m.next[h]]], cfn[m.crnr[m.prev[h]]]
            x0, y0 = math.floor(min(uv1.real, uv2.real)), math.floor(min(uv1.imag, uv2.imag))
            x1, y1 = math.ceil(max(uv1.real, uv2.real)), math.ceil(max(uv1.imag, uv2.imag))
            for x in range(x0, x1 + 1):
                for y in range(y0, y1 + 1):
                    xy = complex(x, y)
                    a = abs(xy - uv1) / abs(uv2 - uv1)
                    if is_collinear(uv1, uv2, xy) and a > EPS and a < 1 - EPS:
                        eq.append((xy, p1 + (p2 - p1) * a, e))
        for f in range(m.nF):
            p = [m.pos[m.idx[f, j]] for j in range(3)]
            uv = [cfn[3 * f + j] for j in range(3)]
            x0 = math.floor(min(u.real for u in uv))
            y0 = math.floor(min(u.imag for u in uv))
            x1 = math.ceil(max(u.real for u in uv))
            y1 = math.ceil(max(u.imag for u in uv))
            for x in range(x0, x1 + 1):
                for y in range(y0, y1 + 1):
                    xy = complex(x, y)
                    if is_inside_triangle(uv[0], uv[1], uv[2], xy):
                        c = calc_coefficient(uv[0], uv[1], uv[2], xy)
                        fq.append((xy, p[0] + (p[1] - p[0]) * c.real + (p[2] - p[0]) * c.imag, f))
        return vq, eq, fq

    def ports(self, cfn, vq, eq, fq):
        """generate_vqvert_qport, generate_eqvert_qport, generate_fqvert_qport in that order."""
        m = self.m
        qp = []

        def link(tmp):
            s = len(tmp)
            for i in range(s):
                tmp[i].idx = len(qp) + i
            for i in range(s):
                tmp[i].prev_id = tmp[(i - 1 + s) % s].idx
                tmp[i].next_id = tmp[(i + 1 + s) % s].idx
            qp.extend(tmp)

        for uvq, pos, v in vq:
            tmp = []
            for h in m.vert_halfs_ccw(v):
                uv1, uv2, uv3 = cfn[m.crnr[m.next[h]]], cfn[m.crnr[m.prev[h]]], cfn[m.crnr[h]]
                r = 0
                while r < 4 and not is_points_into(uv1, uv2, uv3, uv1 + quater_rot(r)):
                    r += 1
                for i in range(4):
                    d = quater_rot((r + i + 3) % 4)
                    f1 = is_points_into(uv1, uv2, uv3, uv1 + d)
                    f2 = is_collinear(uv1, uv2, uv1 + d)
                    f3 = dot(uv2 - uv1, d) > 0
                    if f1 or (f2 and f3):
                        tmp.append(Qport(-1, v, -1, m.face[h], uv1, d, pos))
            link(tmp)
        visit = []
        for uvq, pos, e in eq:
            tmp = []
            h0 = m.edge2half[e]
            for h in (h0, m.twin[h0]):
                c1, c2 = m.crnr[m.next[h]], m.crnr[m.prev[h]]
                uv1, uv2, uv3 = cfn[c1], cfn[c2], cfn[m.crnr[h]]
                p1, p2 = m.pos[self.crnr_vert(c1)], m.pos[self.crnr_vert(c2)]
                uv = lerp(uv1, uv2, math.sqrt(_dot3(pos - p1, pos - p1)) / math.sqrt(_dot3(p2 - p1, p2 - p1)))
                r = 0
                while r < 4 and not is_points_into(uv1, uv2, uv3, uv1 + quater_rot(r)):
                    r += 1
                for i in range(4):
                    d = quater_rot((r + i + 2) % 4)
                    f1 = orientation(uv1, uv2, uv + d) >= 0
                    ev = _normalized(self.conversion_2d_3d(m.face[h], cfn, uv + d))
                    seen = any(math.sqrt(_dot3(ev - w, ev - w)) < EPS for w in visit)
                    if f1 and not seen:
                        tmp.append(Qport(-1, -1, e, m.face[h], nearby_grid1(uv), d, pos))
                        visit.append(ev)
            # the reference asserts 4 ports here (gen_q_port.h:47), compiled out in release
            # builds: the ports found are linked as they are (the visit list is shared by all
            # edge q-vertices, so a probe point another q-vertex already used is skipped)
            link(tmp)
        for uvq, pos, f in fq:
            for i in range(4):
                qp.append(Qport(len(qp), -1, -1, f, uvq, quater_rot(i), pos))
            for i in range(4):
                L = len(qp)
                qp[L - i - 1].next_id = qp[L - (i - 1 + 4) % 4 - 1].idx
                qp[L - i - 1].prev_id = qp[L - (i + 1 + 4) % 4 - 1].idx
        return qp

    def predict_collinear(self, cfn, ori, d, f, pair):
        pb1 = self.conversion_2d_3d(f, cfn, ori)
        pb2 = self.conversion_2d_3d(f, cfn, ori + d)
        pa1 = self.conversion_2d_3d(pair.fid, cfn, pair.uv)
        pa2 = self.conversion_2d_3d(pair.fid, cfn, pair.uv + pair.dir)
        da = _normalized(pa2 - pa1)
        db = _normalized(pb2 - pb1)
        dc = _normalized(pa1 - pb1)
        return abs(1.0 - _dot3(db, dc)) < EPS and abs(1.0 + _dot3(da, db)) < EPS

    def pick_next_half(self, cfn, ori, d, f):
        m = self.m
        for h in m.face_halfs(f):
            uv1, uv2 = cfn[m.crnr[m.next[h]]], cfn[m.crnr[m.prev[h]]]
            ok, rab, rcd = find_strict_intersection(ori, ori + d * 1e2, uv1, uv2)
            if ok and rab > EPS:
                return h, lerp(uv1, uv2, rcd)
        raise RuntimeError("qex: no next half found")

    def edges(self, cfn, matching, qp, max_steps=None):
        """generate_q_edge (qex/gen_q_edge.h:42-119): list of (port idx, port idx)."""
        m = self.m
        heR, heT = self.trs(cfn, matching, 4)
        max_steps = max_steps or 4 * m.nF + 16
        out = []
        for pfr in qp:
            if pfr.connected:
                continue
            ori, d = pfr.uv, pfr.dir
            gri = nearby_grid(ori, d)
            f = pfr.fid
            for _ in range(max_steps):
                if is_inside_triangle(cfn[3 * f], cfn[3 * f + 1], cfn[3 * f + 2], gri):
                    hit = next((p for p in qp if p.idx != pfr.idx and p.fid == f and abs(p.dir - (-d)) < 1e-10
                                and abs(p.uv - gri) < EPS), None)
                    if hit is None:
                        raise RuntimeError(f"qex: port {pfr.idx} finds no face port")
                    pfr.connected = hit.connected = True
                    out.append((pfr.idx, hit.idx))
                    break
                hit = None
                for h in m.face_halfs(f):
                    hit = next((p for p in qp if p.eid >= 0 and not p.connected and p.eid != pfr.eid
                                and p.eid == m.edge[h] and self.predict_collinear(cfn, ori, d, f, p)), None)
                    if hit is not None:
                        break
                if hit is None:
                    for h in m.face_halfs(f):
                        hit = next((p for p in qp if p.vid >= 0 and not p.connected and p.vid != pfr.vid
                                    and p.vid == m.tail[h] and self.predict_collinear(cfn, ori, d, f, p)), None)
                        if hit is not None:
                            break
                if hit is not None:
                    pfr.connected = hit.connected = True
                    out.append((pfr.idx, hit.idx))
                    break
                nh, x = self.pick_next_half(cfn, ori, d, f)
                f = m.face[m.twin[nh]]
                ori = heR[nh] * x + heT[nh]
                d = heR[nh] * d
                gri = nearby_grid(ori, d)
            else:
                raise RuntimeError(f"qex: port {pfr.idx} traced {max_steps} faces without a pair")
        return out

    @staticmethod
    def faces(qp, qe):
        """generate_q_faces (qex/gen_q_face.h): quads as lists of half indices (2 i: edge i
        from port1 to port2, 2 i + 1: reversed)."""
        port1 = []
        port2 = []
        for a, b in qe:
            port1 += [a, b]
            port2 += [b, a]
        first = {}
        for hidx, p in enumerate(port1):
            first.setdefault(p, hidx)
        visit = [False] * len(port1)
        out = []
        start = 0
        while True:
            while start < len(port1) and visit[start]:
                start += 1
            if start == len(port1):
                break
            fh = [start]
            visit[start] = True
            p = port2[start]
            for _ in range(3):
                nxt = first.get(qp[qp[p].prev_id].idx)
                if nxt is None:
                    raise RuntimeError("qex: open quad loop")
                fh.append(nxt)
                visit[nxt] = True
                p = port2[nxt]
            if len(fh) == 4:
                out.append(fh)
        return out


def extract(m, cfn, matching, singular, N=4):
    """example_qex.cpp:73-140: sanitize, then q-verts, ports, edges, faces."""
    q = Qex(m)
    cfn = q.sanitization(matching, singular, N, cfn)
    vq, eq, fq = q.q_verts(cfn)
    qp = q.ports(cfn, vq, eq, fq)
    qe = q.edges(cfn, matching, qp)
    qf = Qex.faces(qp, qe)
    return {"cfn": cfn, "vq": vq, "eq": eq, "fq": fq, "ports": qp, "edges": qe, "faces": qf}
