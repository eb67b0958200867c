"""Pins the batched configuration's Poisson oracle (oracle/poisson.py) on the CPU: the patch
test (a linear function's exact gradient field integrates back to the function), the
stiffness matrix's constant null space and symmetry, and the mesh-set generator's sizes."""
import numpy as np
import pytest

from metriko_amd.meshgen import batch_meshes, icosphere, torus
from oracle.poisson import divergence, pin_first, poisson_uv, stiffness


@pytest.mark.parametrize("name", ["ico3", "torus"])
def test_patch_test_linear_function(name):
    V, F = icosphere(3) if name == "ico3" else torus(16, 24, jitter=0.01, seed=4)
    a = np.array([0.3, -1.2, 0.7])
    u = V @ a
    # exact per-face gradient of the piecewise-linear interpolant of u
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    n = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    g = a[None, :] - (n @ a)[:, None] * n          # tangential part of a = gradient of u on a flat face
    uu, vv = poisson_uv(V, F, g, g)
    assert np.abs(uu - (u - u[0])).max() < 1e-9 * np.abs(u).max()
    assert np.abs(vv - uu).max() == 0.0


def test_stiffness_null_space_and_symmetry():
    V, F = torus(12, 20, jitter=0.02, seed=1)
    L = stiffness(V, F)
    assert abs(L - L.T).max() < 1e-14
    assert np.abs(L @ np.ones(V.shape[0])).max() < 1e-12
    Lp, _ = pin_first(L, np.zeros(V.shape[0]))
    ev = np.linalg.eigvalsh(Lp.toarray())
    assert ev.min() > 0.0


def test_divergence_of_zero_mean_compatibility():
    # the divergence of any field sums to zero (constants are in L's null space)
    V, F = icosphere(2)
    g = np.random.Generator(np.random.PCG64(0)).standard_normal((F.shape[0], 3))
    assert abs(divergence(V, F, g).sum()) < 1e-12


def test_batch_mesh_set():
    ms = batch_meshes()
    nv = [m[0].shape[0] for m in ms]
    assert len(ms) == 64 and min(nv) == 642 and max(nv) <= 160000 and sum(nv) == 1354535
