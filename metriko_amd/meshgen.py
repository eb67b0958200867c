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


def _triple_torus_field(P, seed):
    """Signed field of the three-handle blob: a smooth union (soft-min, k = 16) of three tori
    (R = 0.75, r = 0.3, centres x = -1.4, 0, 1.4 in the z = 0 plane) plus a small random smooth
    bump field drawn from PCG64(seed)."""
    g = np.random.Generator(np.random.PCG64(seed))
    d = []
    for cx in (-1.4, 0.0, 1.4):
        q = np.sqrt((P[..., 0] - cx) ** 2 + P[..., 1] ** 2) - 0.75
        d.append(np.sqrt(q * q + P[..., 2] ** 2) - 0.3)
    k = 16.0
    m = np.minimum(np.minimum(d[0], d[1]), d[2])
    f = m - np.log(sum(np.exp(-k * (di - m)) for di in d)) / k
    for _ in range(6):
        a = 0.012 * g.standard_normal()
        w = 2.5 * g.standard_normal(3)
        ph = g.uniform(0, 2 * np.pi)
        f = f + a * np.sin(P @ w + ph)
    return f


def genus3_blob(h=0.028, seed=4, smooth=20):
    """Closed genus-3 surface: the zero set of the three-handle blob field sampled on a grid of
    spacing h and triangulated by marching tetrahedra (six tetrahedra per cell around the main
    diagonal, so neighbouring cells agree; no ambiguous cases).  Surface points interpolate the
    field linearly along grid edges, clamped to [0.05, 0.95] of the edge; shared edge points
    are welded; triangles face the outside (positive field); then ``smooth`` Taubin iterations.
    h = 0.028 gives about 250k faces."""
    lo = np.array([-2.65, -1.45, -0.70])
    hi = np.array([2.65, 1.45, 0.70])
    n = np.ceil((hi - lo) / h).astype(int) + 1
    xs = [lo[a] + h * np.arange(n[a]) for a in range(3)]
    P = np.stack(np.meshgrid(*xs, indexing="ij"), -1)
    f = _triple_torus_field(P, seed)
    f[f == 0.0] = 1e-12
    nx, ny, nz = n
    gid = lambda i, j, k: (i * ny + j) * nz + k  # noqa: E731
    i, j, k = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    corner = [gid(i, j, k), gid(i + 1, j, k), gid(i + 1, j + 1, k), gid(i, j + 1, k),
              gid(i, j, k + 1), gid(i + 1, j, k + 1), gid(i + 1, j + 1, k + 1), gid(i, j + 1, k + 1)]
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    fv = f.ravel()
    Pv = P.reshape(-1, 3)
    tris, refs = [], []
    for t in tets:
        T = np.stack([corner[c] for c in t], 1)
        inside = fv[T] < 0
        cnt = inside.sum(1)
        for case in (1, 3, 2):
            sel = cnt == case
            if not sel.any():
                continue
            Ts, ins = T[sel], inside[sel]
            if case == 3:
                ins = ~ins  # one outside vertex: same triangle, orientation fixed below
            order = np.argsort(~ins, axis=1, kind="stable")  # marked vertices first
            V4 = np.take_along_axis(Ts, order, 1)
            if case in (1, 3):
                a = V4[:, 0]
                e = [(a, V4[:, 1]), (a, V4[:, 2]), (a, V4[:, 3])]
                tris.append(np.stack([np.stack(x, 1) for x in e], 1))
                refs.append(np.stack([V4[:, 1], a] if case == 1 else [a, V4[:, 1]], 1))
            else:
                a, b, c, d = V4[:, 0], V4[:, 1], V4[:, 2], V4[:, 3]
                q = [(a, c), (a, d), (b, d), (b, c)]
                Q = np.stack([np.stack(x, 1) for x in q], 1)
                tris.append(Q[:, [0, 1, 2]])
                tris.append(Q[:, [0, 2, 3]])
                r = np.stack([c, a], 1)
                refs.append(r)
                refs.append(r)
    E = np.concatenate(tris, 0)                # (m, 3, 2) grid-edge endpoints per triangle corner
    R = np.concatenate(refs, 0)                # (m, 2) (outside, inside) grid vertices
    Es = np.sort(E, axis=2).reshape(-1, 2)
    key = Es[:, 0] * fv.size + Es[:, 1]
    uk, inv = np.unique(key, return_inverse=True)
    a, b = uk // fv.size, uk % fv.size
    t = np.clip(fv[a] / (fv[a] - fv[b]), 0.05, 0.95)
    V = Pv[a] + t[:, None] * (Pv[b] - Pv[a])
    F = inv.reshape(-1, 3).astype(np.int64)
    nrm = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    out = (nrm * (Pv[R[:, 0]] - Pv[R[:, 1]])).sum(1) < 0
    F[out] = F[out][:, [0, 2, 1]]
    return taubin_smooth(V, F, iterations=smooth), F


def taubin_smooth(V, F, iterations=20, lam=0.5, mu=-0.53):
    """Taubin lambda|mu smoothing with uniform neighbour weights (no shrinkage to first
    order): removes the grid-scale ripple of marching tetrahedra, keeps the topology."""
    if iterations <= 0:
        return V
    n = V.shape[0]
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    import scipy.sparse as sp
    A = sp.csr_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
    A = ((A + A.T) > 0).astype(np.float64)
    deg = np.asarray(A.sum(axis=1)).ravel()
    for _ in range(iterations):
        for w in (lam, mu):
            V = V + w * (A @ V / deg[:, None] - V)
    return V
