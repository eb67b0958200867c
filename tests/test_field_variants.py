"""Curvature-aligned field and Principal matching (face_rosy_field.h:25-34,86-98,138-167).

CPU: the oracle's principal-curvature direction against its closed form on a cylinder, its
Principal matching against the Curl matching on smooth edges, and the library's host code
(mk_curvature_dir, mk_field_matching) against the oracle.  GPU: the curvature-aligned solve
(mk_field_solve) against the oracle's sparse direct solve, relative error below 1e-8.
"""
import numpy as np
import pytest

import metriko_amd as mk
from metriko_amd.meshgen import bumpy_icosphere, cylinder
from oracle.field import (FaceRosyField, compute_connection, compute_singular, connection_laplacian,
                          curl_matching, field_roots, principal_curvature_dir, principal_matching,
                          solve_curvature_aligned)
from oracle.hmesh import Hmesh
from tests.helpers import mesh


def test_curvature_dir_cylinder_closed_form():
    """Only axial edges bend (by 2 pi / n, sign + for outward faces); circumferential and
    diagonal edges join coplanar triangles.  So every face has D = -(dz alpha / 4) e^{2 i phi_z},
    phi_z the angle of the axis in the face basis."""
    n, rows, dz = 24, 6, 0.13
    V, F = cylinder(n, rows, r=1.0, dz=dz)
    m = Hmesh(V, F)
    D = principal_curvature_dir(m)
    ez = m.faceBasisX[:, 2] + 1j * m.faceBasisY[:, 2]
    expect = -(dz * 2 * np.pi / n / 4) * ez ** 2
    assert np.abs(D - expect).max() < 1e-13


def test_curvature_aligned_solve_residual():
    """The oracle's solve: L x = M D^2 / sqrt(D^H M D) (N = 4) and D / sqrt(...) (N = 2)."""
    V, F = mesh("torus_small")
    m = Hmesh(V, F)
    D = principal_curvature_dir(m)
    for N, Dn in ((2, D), (4, D * D)):
        L = connection_laplacian(m, N)
        x = solve_curvature_aligned(L, m.faceArea, D, N)
        rhs = m.faceArea * Dn / np.sqrt(np.sum(m.faceArea * np.abs(Dn) ** 2))
        assert np.linalg.norm(L @ x - rhs) <= 1e-10 * np.linalg.norm(rhs)


def test_principal_matches_curl_on_smooth_edges():
    """Where the transported roots nearly coincide both matchings pick the same shift, and the
    singularities (from the matching-independent effort) agree everywhere."""
    V, F = mesh("icosphere_obj")
    m = Hmesh(V, F)
    f = FaceRosyField(m, 4)
    conn = compute_connection(m)
    mc, ec = curl_matching(m, f.field, conn, 4)
    mp, ep = principal_matching(m, f.field, conn, 4)
    smooth = np.abs(ep) < 0.3
    assert smooth.sum() > 0.9 * len(ep)
    assert np.array_equal(np.mod(mp[smooth], 4), mc[smooth])
    assert np.array_equal(compute_singular(m, ec, 4), compute_singular(m, ep, 4))


@pytest.mark.parametrize("name", ["icosphere_obj", "torus_small"])
def test_host_curvature_dir_matches_oracle(name):
    V, F = mesh(name)
    D = mk.curvature_dir(V, F)
    ref = principal_curvature_dir(Hmesh(V, F))
    assert np.abs(D - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("kind", ["principal", "curl"])
@pytest.mark.parametrize("name", ["icosphere_obj", "torus_small"])
def test_host_matching_matches_oracle(name, kind):
    V, F = mesh(name)
    m = Hmesh(V, F)
    f = FaceRosyField(m, 4)
    out = mk.field_matching(V, F, f.compressed, 4, kind)
    conn = compute_connection(m)
    fn = principal_matching if kind == "principal" else curl_matching
    mat, eff = fn(m, field_roots(f.compressed, 4), conn, 4)
    assert np.array_equal(out["matching"], mat)
    assert np.array_equal(out["singular"], compute_singular(m, eff, 4))
    assert np.abs(out["effort"] - eff).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("N", [2, 4])
@pytest.mark.parametrize("name", ["icosphere_obj", "torus_small", "bumpy3"])
def test_gpu_curvature_aligned_matches_oracle(gpu_ctx, name, N):
    V, F = bumpy_icosphere(3, jitter=2e-4, seed=3) if name == "bumpy3" else mesh(name)
    m = Hmesh(V, F)
    x, st = gpu_ctx.field(V, F, N, kind="curvature")
    ref = solve_curvature_aligned(connection_laplacian(m, N), m.faceArea, principal_curvature_dir(m), N)
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    print(name, N, "curvature-aligned rel err", err, st)
    assert err < 1e-8
    assert st["rel_residual"] <= 1e-10
