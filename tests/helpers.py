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
