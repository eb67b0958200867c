"""Host sparse kernels of include/metriko/core/common/sparse.h (the multigrid and IGM set-up):
the transpose's 32- and 64-bit counter paths and its chunked OpenMP form against a one-thread
counting transpose, the complex product At * A against ordered sums of std::complex products
(bit for bit, rows on both sides of the sort-path threshold), and the inline complex product
against std::complex's own (tests/native/sparse_host.cpp, compiled here with g++ -fopenmp).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def sparse_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("native") / "sparse_host"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "sparse_host.cpp"), "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("threads", [1, 4])
def test_transpose_and_product_match_ordered_restatements(sparse_exe, seed, threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sparse_exe, str(seed)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
