"""Pins of the seam restatement (face_rosy_field.h:268-360 with libigl v2.5.0's
cut_mesh_from_singularities and dijkstra, oracle/seam.py).

* Hand-worked octahedron cases: the genus-0 spanning-tree cut retracts to nothing, then one
  unit-weight Dijkstra per singular vertex (ascending ids) to the vertices already on the cut,
  the queue ordered by (distance, vertex) and predecessors set on strict improvement only.
* The MeshCutter restatement against its characterisation, computed with networkx: the
  sides left cut by a breadth-first search over the dual graph from face 0 (neighbours in
  side order 0, 1, 2, faces marked when queued), reduced to the 2-core of the primal graph
  they form (Retract dissolves every cut edge with an endpoint of cut degree 1 until none is
  left).  On a 4 x 6 torus written out face by face and on finer tori: two loops.
* Dijkstra against its characterisation: the target reached is the one with the smallest
  (BFS distance, id); each vertex's predecessor is its smallest-id neighbour one step closer.
"""
import networkx as nx
import numpy as np
import pytest

from oracle.seam import cut_mesh_from_singularities_zero_matching, cut_mesh_with_singularities, dijkstra_unit, \
    adjacency_list

# octahedron: 0 = +x, 1 = -x, 2 = +y, 3 = -y, 4 = +z, 5 = -z; outward counter-clockwise faces
OCTA_F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])


def _cut_edges(F, cuts):
    return sorted({tuple(sorted((int(F[f, s]), int(F[f, (s + 1) % 3])))) for f, s in zip(*np.nonzero(cuts))})


def test_octahedron_spanning_cut_is_empty():
    cuts, _, _ = cut_mesh_from_singularities_zero_matching(OCTA_F, 6)
    assert not cuts.any()


def test_octahedron_singulars_1_2_5():
    """s = 1: nothing on the cut yet, the path is [1].  s = 2: 2's neighbours 0, 1, 4, 5 get
    distance 1; (1, 0) pops before (1, 1), 0 is no target, then 1 is: path [1, 2], cut on the
    first common face of 1 and 2 (faces {1, 5}: face 1 = (2, 1, 4), side 0) and its twin (face 5,
    side 0).  s = 5: neighbours 0, 1, 2, 3 at distance 1; 0 pops first (no target), then 1:
    path [1, 5] on face 5 = (1, 2, 5), side 2, and face 6 = (3, 1, 5), side 1."""
    cuts = cut_mesh_with_singularities(OCTA_F, 6, [1, 2, 5])
    assert _cut_edges(OCTA_F, cuts) == [(1, 2), (1, 5)]
    assert sorted(zip(*np.nonzero(cuts))) == [(1, 0), (5, 0), (5, 2), (6, 1)]


def test_octahedron_singulars_0_1():
    """s = 1 to the cut {0}: all four shortest paths 1-x-0 (x = 2, 3, 4, 5) tie; 2 pops first
    at distance 1 and sets prev[0] = 2; 3, 4, 5 reach 0 at distance 2 too but do not improve
    it.  Path [0, 2, 1]: edge 0-2 on face 0 side 0 (twin face 4 side 0), edge 2-1 on face 1
    side 0 (twin face 5 side 0)."""
    cuts = cut_mesh_with_singularities(OCTA_F, 6, [0, 1])
    assert _cut_edges(OCTA_F, cuts) == [(0, 2), (1, 2)]
    assert sorted(zip(*np.nonzero(cuts))) == [(0, 0), (1, 0), (4, 0), (5, 0)]


def grid_torus(n, m):
    """Vertex v(i, j) = i m + j on an n x m periodic grid; cell (i, j) gives faces
    2(i m + j) = (a, b, c) and 2(i m + j) + 1 = (a, c, d) with a = v(i, j), b = v(i, j+1),
    c = v(i+1, j+1), d = v(i+1, j)."""
    F = []
    for i in range(n):
        for j in range(m):
            a, b = i * m + j, i * m + (j + 1) % m
            c, d = ((i + 1) % n) * m + (j + 1) % m, ((i + 1) % n) * m + j
            F += [(a, b, c), (a, c, d)]
    return np.array(F)


def _meshcutter_characterisation(F):
    """2-core of the primal edges not crossed by the BFS dual tree from face 0 (networkx)."""
    nF = F.shape[0]
    side_of = {}
    for f in range(nF):
        for s in range(3):
            side_of[(min(F[f, s], F[f, (s + 1) % 3]), max(F[f, s], F[f, (s + 1) % 3]), f)] = s
    owners = {}
    for (a, b, f) in side_of:
        owners.setdefault((a, b), []).append(f)
    D = nx.DiGraph()
    D.add_nodes_from(range(nF))
    for f in range(nF):
        for s in range(3):
            e = (min(F[f, s], F[f, (s + 1) % 3]), max(F[f, s], F[f, (s + 1) % 3]))
            g = [x for x in owners[e] if x != f]
            if g:
                D.add_edge(f, g[0])
    crossed = set()
    for f, g in nx.bfs_edges(D, 0):
        crossed.add(tuple(sorted(set(map(int, F[f])) & set(map(int, F[g])))))
    P = nx.Graph()
    P.add_edges_from(e for e in owners if e not in crossed)
    return sorted(tuple(map(int, e)) for e in nx.k_core(P, 2).edges()) if P.number_of_edges() else []


@pytest.mark.parametrize("nm", [(4, 6), (5, 7), (16, 24)])
def test_meshcutter_is_two_core_of_bfs_cotree(nm):
    F = grid_torus(*nm)
    cuts, _, _ = cut_mesh_from_singularities_zero_matching(F, nm[0] * nm[1])
    got = _cut_edges(F, cuts)
    assert got == sorted(tuple(sorted(e)) for e in _meshcutter_characterisation(F))
    # genus 1: the cut graph has V - E = 1 - 2g = -1 (two loops through one vertex or joined)
    verts = {v for e in got for v in e}
    assert len(verts) - len(got) == -1


def test_meshcutter_4x6_torus_known_cut():
    """The 4 x 6 torus of grid_torus: the cut is a meridian and a longitude (two loops)."""
    F = grid_torus(4, 6)
    cuts, _, _ = cut_mesh_from_singularities_zero_matching(F, 24)
    got = _cut_edges(F, cuts)
    G = nx.Graph(got)
    assert nx.number_connected_components(G) == 1
    assert len(nx.cycle_basis(G)) == 2
    assert all(d >= 2 for _, d in G.degree())


@pytest.mark.parametrize("seed", range(6))
def test_dijkstra_characterisation(seed):
    F = grid_torus(6, 9)
    nV = 54
    VV = adjacency_list(F, nV)
    g = np.random.default_rng(seed)
    src = int(g.integers(nV))
    targets = set(int(t) for t in g.choice(nV, size=4, replace=False)) - {src}
    found, prev = dijkstra_unit(src, targets, VV)
    G = nx.Graph()
    for v in range(nV):
        G.add_edges_from((v, w) for w in VV[v])
    dist = nx.single_source_shortest_path_length(G, src)
    assert found == min(targets, key=lambda t: (dist[t], t))
    v = found
    while v != src:   # the predecessor chain: smallest-id neighbour one step closer
        p = min(w for w in VV[v] if dist[w] == dist[v] - 1)
        assert prev[v] == p
        v = p
