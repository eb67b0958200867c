"""Host unit test of the captured-graph keys: a rebuilt hierarchy or another context at a
reused address never replays a graph captured through the old one (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_graph_key_changes_when_hierarchy_rebuilt(tmp_path):
    exe = tmp_path / "graph_key_test"
    src = os.path.join(ROOT, "tests", "native", "graph_key_test.cpp")
    inc = os.path.join(ROOT, "metriko_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", inc, src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
