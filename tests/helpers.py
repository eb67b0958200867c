"""Shared helpers for the parity tests (meshes, phase-aligned field error)."""
import os

import numpy as np

from metriko_amd.meshgen import icosphere, read_obj, torus

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mesh(name):
    if name == "ico2":
        return icosphere(2)
    if name == "ico3":
        return icosphere(3)
    if name == "icosphere_obj":
        return read_obj(os.path.join(GOLDEN, "icosphere.obj"))
    if name == "torus_small":
        return torus(16, 24, R=1.0, r=0.35, jitter=0.002, seed=2)
    if name == "torus_mid":
        return torus(32, 48, R=1.0, r=0.35, jitter=0.002, seed=2)
    raise KeyError(name)


def phase_error(x, ref):
    """min over |c| = 1 of ||x - c ref|| / ||ref||."""
    c = np.vdot(ref, x)
    c = c / abs(c) if abs(c) > 0 else 1.0
    return np.linalg.norm(x - c * ref) / np.linalg.norm(ref)


def torus_config():
    """The 204,800-face torus of the FP32-vs-FP64 configuration: R = 1, r = 0.35 on a
    256 x 400 grid, N(0, 0.002) jitter along the normals (seed 1); u0 from PCG64(1)."""
    from metriko_amd.meshgen import torus
    V, F = torus(256, 400, 1.0, 0.35, jitter=0.002, seed=1)
    g = np.random.Generator(np.random.PCG64(1))
    u0 = g.uniform(-1, 1, F.shape[0]) + 1j * g.uniform(-1, 1, F.shape[0])
    return V, F, u0


def cut_hash(idx):
    """Order-sensitive checksum of the cut-mesh corner indices."""
    w = (np.arange(idx.size, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(1)) & np.uint64((1 << 40) - 1)
    return int((w * np.asarray(idx).reshape(-1).astype(np.uint64)).sum() & np.uint64((1 << 61) - 1))
