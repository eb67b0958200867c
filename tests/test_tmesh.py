"""QGP T-mesh (include/metriko/core/tmesh/tmesh.h through mk_tmesh_build, host code) against
the oracle (oracle/tmesh.py) on the motorcycle graphs of the injective naive uv, plus what the
construction fixes: every half-edge lies in exactly one T-quad, every T-quad goes round its
four sides 0..3 once (cyclically non-decreasing, four side changes, not starting inside a
side), each curve contributes one half-edge leaving its singular vertex, and
the T-mesh of a sphere has Euler characteristic 2 (vertices: the singular vertices with
ports and one per crash)."""
import numpy as np
import pytest

import metriko_amd as mk
from metriko_amd.meshgen import icosphere, bumpy_icosphere
from oracle.motorcycle import MotorcycleGraph
from oracle.pipeline import run_pipeline
from oracle.tmesh import Tmesh


@pytest.mark.parametrize("mesh,ratio", [("ico2", 0.1), ("ico3", 0.05), ("bumpy3", 0.1)])
def test_tmesh(mesh, ratio):
    V, F = {"ico2": lambda: icosphere(2), "ico3": lambda: icosphere(3), "bumpy3": lambda: bumpy_icosphere(3)}[mesh]()
    r = run_pipeline(V, F, ratio=ratio, seamless=False, local_injectivity=True)
    uv = r["uv"]
    mt = np.asarray(r["combed_matching"], np.int32)
    sg = np.asarray(r["singular"], np.int32)
    g = MotorcycleGraph(r["mesh"], uv[:, 0] + 1j * uv[:, 1], mt, sg)
    ref = Tmesh(g)
    got = mk.tmesh_build(V, F, mt, sg, uv)
    assert np.array_equal(got["tedges"], np.array(ref.tedges, np.int32))
    assert got["tquads"] == [(list(t), list(s)) for t, s in ref.tquads]
    assert got["th2sing"].tolist() == ref.th2sing
    # invariants
    nth = 2 * len(got["tedges"])
    cover = sorted(t for q in got["tquads"] for t in q[0])
    assert cover == list(range(nth))
    for thids, sides in got["tquads"]:
        steps = [(b - a) % 4 for a, b in zip(sides, sides[1:] + sides[:1])]
        assert set(steps) <= {0, 1} and steps.count(1) == 4 and sides[0] != sides[-1]
    ncurves = len(g.ports)
    assert sum(1 for x in got["th2sing"] if x >= 0) == ncurves
    nv = len(set(p["vert"] for p in g.ports)) + ncurves
    assert nv - len(got["tedges"]) + len(got["tquads"]) == 2
