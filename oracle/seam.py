"""Oracle: seam graph, cut mesh, combing, extrinsic field (CPU combinatorics).

Restates ``vectorfield/face_rosy_field.h:207-360`` and ``igm/parameterization.h:126-152``,
plus the libigl v2.5.0 routines they call (``triangle_triangle_adjacency``,
``vertex_triangle_adjacency``, ``adjacency_list``, ``cut_mesh_from_singularities``,
``dijkstra``), restated from libigl's published sources.  Test infrastructure
only; see ``oracle/__init__.py``.
"""
import bisect
import heapq
from collections import deque

import numpy as np

from .hmesh import Hmesh


def triangle_triangle_adjacency(F):
    """libigl ``triangle_triangle_adjacency(F, TT, TTi)``: TT(f,i) is the face across
    side i (F(f,i)->F(f,(i+1)%3)), TTi(f,i) its side index there; -1 on borders."""
    nF = F.shape[0]
    key = {}
    for f in range(nF):
        for i in range(3):
            a, b = int(F[f, i]), int(F[f, (i + 1) % 3])
            key.setdefault((min(a, b), max(a, b)), []).append((f, i))
    TT = -np.ones((nF, 3), dtype=np.int64)
    TTi = -np.ones((nF, 3), dtype=np.int64)
    for lst in key.values():
        if len(lst) == 2:
            (f, i), (g, j) = lst
            TT[f, i], TTi[f, i] = g, j
            TT[g, j], TTi[g, j] = f, i
    return TT, TTi


def vertex_triangle_adjacency(F, nV):
    """libigl ``vertex_triangle_adjacency``: VF[v] = faces containing v, in face order."""
    VF = [[] for _ in range(nV)]
    for f in range(F.shape[0]):
        for j in range(3):
            VF[int(F[f, j])].append(f)
    return VF


def adjacency_list(F, nV):
    """libigl ``adjacency_list(F, A)``: neighbours per vertex, sorted ascending and deduplicated."""
    A = [set() for _ in range(nV)]
    for f in range(F.shape[0]):
        for j in range(3):
            s, d = int(F[f, j]), int(F[f, (j + 1) % 3])
            A[s].add(d)
            A[d].add(s)
    return [sorted(a) for a in A]


def cut_mesh_from_singularities_zero_matching(F, nV):
    """libigl ``cut_mesh_from_singularities(V, F, MMatch=0, cuts)`` (``MeshCutter``).

    Called with an all-zero matching at face_rosy_field.h:279, so no side is a
    rotation seam.  Algorithm (libigl v2.5.0):
      1. all sides start cut; for each unvisited face in index order, a BFS
         (deque, sides 0,1,2) un-cuts every side crossed into an unvisited face;
      2. ``Retract``: per-vertex count of cut edges (each edge once: the side
         with TT(f,s) <= f); then passes over (f, s) in order dissolve a cut
         side whose first vertex F(f,s) has count 1, until a pass changes
         nothing or the pass guard (>10000) trips.
    """
    nF = F.shape[0]
    TT, TTi = triangle_triangle_adjacency(F)
    seams = np.ones((nF, 3), dtype=bool)
    visited = np.zeros(nF, dtype=bool)
    for f0 in range(nF):
        if visited[f0]:
            continue
        visited[f0] = True
        d = deque([f0])
        while d:
            f = d.popleft()
            for s in range(3):
                g, j = TT[f, s], TTi[f, s]
                if j == -1:
                    g, j = f, s
                if not visited[g]:
                    seams[f, s] = False
                    seams[g, j] = False
                    visited[g] = True
                    d.append(g)
    e = np.zeros(nV, dtype=np.int64)
    for f in range(nF):
        for s in range(3):
            if seams[f, s] and TT[f, s] <= f:
                e[F[f, s]] += 1
                e[F[f, (s + 1) % 3]] += 1
    guard = 0
    while True:
        over = True
        for f in range(nF):
            for s in range(3):
                if seams[f, s] and e[F[f, s]] == 1:
                    seams[f, s] = False
                    if TT[f, s] != -1:
                        seams[TT[f, s], TTi[f, s]] = False
                    e[F[f, s]] -= 1
                    e[F[f, (s + 1) % 3]] -= 1
                    over = False
        guard += 1
        if guard > 10001:
            over = True
        if over:
            break
    return seams, TT, TTi


def dijkstra_unit(source, targets, VV):
    """libigl ``dijkstra(source, targets, VV, min_distance, previous)`` with unit weights.

    Queue is a std::set of (distance, vertex) pairs; a vertex is relaxed only on
    a strict improvement; returns the first target popped, or -1.
    """
    nV = len(VV)
    dist = np.full(nV, np.inf)
    prev = -np.ones(nV, dtype=np.int64)
    dist[source] = 0.0
    heap = [(0.0, source)]
    done = np.zeros(nV, dtype=bool)
    while heap:
        du, u = heapq.heappop(heap)
        if done[u] or du != dist[u]:
            continue
        done[u] = True
        if u in targets:
            return u, prev
        for v in VV[u]:
            nd = du + 1.0
            if nd < dist[v]:
                dist[v] = nd
                prev[v] = u
                heapq.heappush(heap, (nd, v))
    return -1, prev


def cut_mesh_with_singularities(F, nV, singularities):
    """metriko ``cut_mesh_with_singularities`` (face_rosy_field.h:268-343)."""
    cuts, TT, TTi = cut_mesh_from_singularities_zero_matching(F, nV)
    VF = vertex_triangle_adjacency(F, nV)
    VV = adjacency_list(F, nV)
    in_cut = set()
    for i in range(F.shape[0]):
        for j in range(3):
            if cuts[i, j]:
                in_cut.add(int(F[i, j]))
    for s in singularities:
        s = int(s)
        found, prev = dijkstra_unit(s, in_cut, VV)
        if found == -1:
            path = [s]
        else:
            path = []
            v = found
            while v != -1:
                path.append(int(v))
                v = prev[v]
        in_cut.update(path)
        for ii in range(len(path) - 1):
            v0, v1 = path[ii], path[ii + 1]
            common = sorted(set(VF[v0]) & set(VF[v1]))
            assert len(common) == 2
            fi = common[0]
            j = -1
            for z in range(3):
                a, b = F[fi, z], F[fi, (z + 1) % 3]
                if (a == v0 and b == v1) or (a == v1 and b == v0):
                    j = z
            assert j != -1
            cuts[fi, j] = True
            cuts[TT[fi, j], TTi[fi, j]] = True
    return cuts


def compute_seam(m, singular):
    """``compute_seam`` (face_rosy_field.h:345-360): singular vertices (value != 0)
    in ascending id, cut, then seam[face2edge(f,j)] for each cut side."""
    svids = [v for v in range(m.nV) if singular[v] != 0]
    cuts = cut_mesh_with_singularities(m.idx, m.nV, svids)
    seam = np.zeros(m.nE, dtype=bool)
    seam[m.face2edge[cuts]] = True
    return seam


def compute_cut_mesh(m, seam):
    """``compute_cut_mesh`` (parameterization.h:126-152).

    Per vertex: start at the first seam (or boundary) halfedge found clockwise
    from v.half(), walk counter-clockwise and open a new cut vertex at the start
    and at every seam crossing.
    """
    cut2pos = []
    cutF = -np.ones((m.nF, 3), dtype=np.int64)
    for v in range(m.nV):
        bgn = m.vert2half[v]
        if m.isBV[v]:
            for h in m.vert_halfs_cw(v):
                if m.edge_is_boundary[m.edge[h]]:
                    bgn = h
                    break
        else:
            for h in m.vert_halfs_cw(v):
                if seam[m.edge[h]]:
                    bgn = h
                    break
        for h in m.vert_halfs_ccw(v, bgn):
            if m.face[h] == -1:
                continue
            if seam[m.edge[h]] or h == bgn:
                cut2pos.append(m.pos[v])
            f = m.face[h]
            for j in range(3):
                if m.idx[f, j] == v:
                    cutF[f, j] = len(cut2pos) - 1
    cutV = np.array(cut2pos)
    return Hmesh(cutV, cutF)


def compute_combed_field(m, field, matching, seam, N):
    """``compute_combbed_field`` (face_rosy_field.h:207-249): BFS from face 0 over
    non-seam interior edges, rotating the root labels by the accumulated turns;
    returns (combed field, new matching)."""
    nF = m.nF
    out = np.zeros_like(field)
    turns = np.zeros(nF, dtype=np.int64)
    visit = np.zeros(nF, dtype=bool)
    q = deque([(0, 0)])
    while q:
        i, mm = q.popleft()
        r = N - mm
        if visit[i]:
            continue
        visit[i] = True
        turns[i] = mm
        out[i, 0:r] = field[i, mm:N]
        out[i, r:N] = field[i, 0:mm]
        for h in m.face_halfs(i):
            e = m.edge[h]
            b = m.edge2face[e, 0] == i
            fn = m.edge2face[e, 1] if b else m.edge2face[e, 0]
            n = (1 if b else -1) * matching[e]
            n = (n + mm + 1000 * N) % N
            if (not m.edge_is_boundary[e]) and (not visit[fn]) and (not seam[e]):
                q.append((fn, n))
    new_matching = -np.ones(m.nE, dtype=np.int64)
    e = np.nonzero(~m.edge_is_boundary)[0]
    new_matching[e] = (turns[m.edge2face[e, 0]] - turns[m.edge2face[e, 1]] + matching[e] + 1000 * N) % N
    return out, new_matching


def compute_extrinsic_field(m, field):
    """``compute_extrinsic_field`` (face_rosy_field.h:251-264): nF x 12, four unit 3-vectors."""
    ext = np.zeros((m.nF, 12))
    for j in range(4):
        c = field[:, j]
        w = c.real[:, None] * m.faceBasisX + c.imag[:, None] * m.faceBasisY
        ext[:, 3 * j:3 * j + 3] = w / np.sqrt((w * w).sum(axis=1))[:, None]
    return ext
