"""The file formats of the example drivers: the triangle OBJ reader (igl::readOBJ as
example_igm.cpp / example_qex.cpp use it) and the corner-uv JSON dump of
example_qex.cpp:14-43, through the library (host code)."""
import json

import numpy as np
import pytest

import metriko_amd as mk
from metriko_amd.meshgen import read_obj as read_obj_py

OBJ = """# two triangles
mtllib x.mtl
v 0 0 0
v 1.5 0 0
vt 0 0
vn 0 0 1
v 1 1 0.25
v -1e-3 1 0
f 1 2 3
f 1/1/1 3//1 4/2
"""


def test_obj_reader(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text(OBJ)
    V, F = mk.read_obj(p)
    np.testing.assert_array_equal(V, [[0, 0, 0], [1.5, 0, 0], [1, 1, 0.25], [-1e-3, 1, 0]])
    np.testing.assert_array_equal(F, [[0, 1, 2], [0, 2, 3]])
    Vp, Fp = read_obj_py(p)
    np.testing.assert_array_equal(V, Vp)
    np.testing.assert_array_equal(F, Fp)


def test_obj_reader_rejects_polygons(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    with pytest.raises(mk.MetrikoError):
        mk.read_obj(p)


def test_cfn_json_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    c = rng.standard_normal(30) * 1e3 + 1j * rng.standard_normal(30) * 1e-7
    p = tmp_path / "uv.json"
    mk.write_cfn_json(p, c)
    d = json.loads(p.read_text())
    assert list(d) == ["cfn"] and len(d["cfn"]) == 60
    np.testing.assert_array_equal(np.array(d["cfn"][0::2]) + 1j * np.array(d["cfn"][1::2]), c)
    np.testing.assert_array_equal(mk.read_cfn_json(p), c)
    # a dump written by another JSON writer (spaces, exponents) reads the same
    p.write_text(json.dumps({"cfn": [1.0, -2.5e-3, 3, 4]}, indent=2))
    np.testing.assert_array_equal(mk.read_cfn_json(p), [1 - 2.5e-3j, 3 + 4j])
