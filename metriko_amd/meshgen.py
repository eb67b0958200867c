"""Synthetic, seeded input meshes for the parity tests and the benchmark.

All randomness uses numpy's ``Generator(PCG64(seed))`` (bit-stable across
platforms); jitter is applied along the analytic surface normal.
"""
import numpy as np


def read_obj(path):
    """Minimal Wavefront OBJ reader for ``v x y z`` and triangle ``f a b c`` lines
    (1-based, ``a/b/c`` forms accepted), as ``igl::readOBJ`` returns them for the
    reference's ``example/models/icosphere.obj``."""
    V, F = [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("v "):
                V.append([float(t) for t in line.split()[1:4]])
            elif line.startswith("f "):
                F.append([int(t.split("/")[0]) - 1 for t in line.split()[1:4]])
    return np.array(V, dtype=np.float64), np.array(F, dtype=np.int64)


def icosphere(level, radius=1.0):
    """Procedural icosphere: icosahedron subdivided ``level`` times (1 -> 4 triangles,
    midpoints projected to the sphere).  New vertices are numbered in the sorted
    order of their parent edges (min, max).  nV = 10*4^L + 2, nF = 20*4^L."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    V /= np.linalg.norm(V, axis=1)[:, None]
    for _ in range(level):
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        E = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1).reshape(-1, 2)
        Es = np.sort(E, axis=1)
        key = Es[:, 0] * (V.shape[0] + 1) + Es[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        first = np.zeros(len(uk), dtype=np.int64)
        first[inv[::-1]] = np.arange(len(inv))[::-1]
        ue = Es[first]
        mid = 0.5 * (V[ue[:, 0]] + V[ue[:, 1]])
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        base = V.shape[0]
        m = (base + inv).reshape(-1, 3)
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        F = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 0)
        V = np.concatenate([V, mid], 0)
    return V * radius, F


def torus(n_tube, n_ring, R=1.0, r=0.35, jitter=0.0, seed=0):
    """Torus on an n_ring x n_tube grid: phi_i = 2 pi i / n_ring around the main
    circle, theta_j = 2 pi j / n_tube around the tube; vertex (i, j) has id
    i*n_tube + j; each quad (i,j),(i+1,j),(i+1,j+1),(i,j+1) is split on its
    (i,j)-(i+1,j+1) diagonal; outward orientation; optional N(0, jitter)
    displacement along the analytic normal with seed ``seed``."""
    i = np.arange(n_ring)[:, None]
    j = np.arange(n_tube)[None, :]
    phi = 2 * np.pi * i / n_ring
    th = 2 * np.pi * j / n_tube
    x = (R + r * np.cos(th)) * np.cos(phi)
    y = (R + r * np.cos(th)) * np.sin(phi)
    z = r * np.sin(th) * np.ones_like(phi)
    V = np.stack([x, y, z], -1).reshape(-1, 3)
    nrm = np.stack([np.cos(th) * np.cos(phi), np.cos(th) * np.sin(phi), np.sin(th) * np.ones_like(phi)],
                   -1).reshape(-1, 3)
    if jitter > 0:
        g = np.random.Generator(np.random.PCG64(seed))
        V = V + nrm * (jitter * g.standard_normal(V.shape[0]))[:, None]
    ii, jj = np.meshgrid(np.arange(n_ring), np.arange(n_tube), indexing="ij")
    a = ii * n_tube + jj
    b = ((ii + 1) % n_ring) * n_tube + jj
    c = ((ii + 1) % n_ring) * n_tube + (jj + 1) % n_tube
    d = ii * n_tube + (jj + 1) % n_tube
    F = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)], 0)
    # interleave the two triangles of each quad for locality
    nq = n_ring * n_tube
    F = np.stack([F[:nq], F[nq:]], 1).reshape(-1, 3)
    return V, F.astype(np.int64)


def cylinder(n_around, n_rows, r=1.0, dz=0.1):
    """Open cylinder: n_rows + 1 rings of n_around vertices, ring j at z = j dz, vertex (i, j)
    at angle 2 pi i / n_around with id j*n_around + i; quad (i,j),(i+1,j),(i+1,j+1),(i,j+1)
    split on its (i,j)-(i+1,j+1) diagonal; outward orientation."""
    j, i = np.meshgrid(np.arange(n_rows + 1), np.arange(n_around), indexing="ij")
    th = 2 * np.pi * i / n_around
    V = np.stack([r * np.cos(th), r * np.sin(th), dz * j], -1).reshape(-1, 3).astype(np.float64)
    jj, ii = np.meshgrid(np.arange(n_rows), np.arange(n_around), indexing="ij")
    a = jj * n_around + ii
    b = jj * n_around + (ii + 1) % n_around
    c = (jj + 1) * n_around + (ii + 1) % n_around
    d = (jj + 1) * n_around + ii
    F = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], -2).reshape(-1, 3)
    return V, F.astype(np.int64)


def bumpy_icosphere(level, amp=0.05, jitter=1e-3, seed=5):
    """Icosphere with radial displacement r = 1 + amp*sin(6 theta)*sin(5 phi)
    (theta = acos z, phi = atan2(y, x)) plus N(0, jitter) radial jitter, seed ``seed``."""
    V, F = icosphere(level)
    th = np.arccos(np.clip(V[:, 2], -1, 1))
    ph = np.arctan2(V[:, 1], V[:, 0])
    rr = 1.0 + amp * np.sin(6 * th) * np.sin(5 * ph)
    if jitter > 0:
        g = np.random.Generator(np.random.PCG64(seed))
        rr = rr + jitter * g.standard_normal(V.shape[0])
    return V * rr[:, None], F


def jittered(V, F, sigma, seed):
    """Displace vertices by N(0, sigma) along area-weighted vertex normals."""
    e1 = V[F[:, 1]] - V[F[:, 0]]
    e2 = V[F[:, 2]] - V[F[:, 0]]
    fn = np.cross(e1, e2)
    vn = np.zeros_like(V)
    for k in range(3):
        np.add.at(vn, F[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1)[:, None]
    g = np.random.Generator(np.random.PCG64(seed))
    return V + vn * (sigma * g.standard_normal(V.shape[0]))[:, None], F


def batch_meshes(count=64, seed=3, max_torus_verts=160000, max_level=6):
    """The batched set: ``count`` closed meshes of uneven size drawn from PCG64(seed).  Each
    mesh is an icosphere of level 3..max_level (642 / 2,562 / 10,242 / 40,962 v) or a torus
    (R = 1, r = 0.35) whose vertex count is log-uniform in [2,000, max_torus_verts] on an
    n_ring x n_tube grid with n_ring / n_tube = 1.5625 (the 400 x 256 aspect); every mesh gets
    N(0, 0.002) jitter along its normal (its own seed).  Returns a list of (V, F)."""
    g = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(count):
        sub = 1000 * seed + i
        if g.random() < 0.5:
            level = int(g.integers(3, max_level + 1))
            V, F = icosphere(level)
            h = np.random.Generator(np.random.PCG64(sub))
            V = V * (1.0 + 0.002 * h.standard_normal(V.shape[0]))[:, None]
        else:
            nv = float(np.exp(g.uniform(np.log(2000.0), np.log(max_torus_verts))))
            n_tube = max(8, int(round((nv / 1.5625) ** 0.5)))
            n_ring = max(8, int(round(1.5625 * n_tube)))
            V, F = torus(n_tube, n_ring, 1.0, 0.35, jitter=0.002, seed=sub)
        out.append((V, F))
    return out
