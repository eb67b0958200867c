"""Oracle: the example_igm driver (``example/example_igm.cpp:14-36``) as one call.

Test infrastructure only; see ``oracle/__init__.py``.
"""
import time

import numpy as np

from . import progress
from .field import FaceRosyField
from .hmesh import Hmesh
from .igm import RosyParameterization
from .seam import compute_combed_field, compute_cut_mesh, compute_extrinsic_field, compute_seam


def run_pipeline(V, F, N=4, ratio=0.03, u0=None, n_iter=50, literal=False, seamless=True,
                 round_seams=False, local_injectivity=False):
    """Hmesh -> Smoothest field -> Curl matching -> seam -> cut -> comb -> IGM.

    Flags as ``example_igm.cpp:24-27`` (seamless, no injectivity, roundSeams=false);
    ``ratio`` is the RosyParameterization grid scale (parameterization.h:51).
    Returns a dict of every intermediate the HIP path is compared against.
    """
    t = {}
    t0 = time.perf_counter()
    mesh = Hmesh(V, F)
    t["hmesh"] = time.perf_counter() - t0
    progress.report("hmesh built")
    t0 = time.perf_counter()
    rawf = FaceRosyField(mesh, N, u0=u0, n_iter=n_iter)
    t["field"] = time.perf_counter() - t0
    rawf.compute_matching("curl")
    progress.report("matching and singularities done")
    t0 = time.perf_counter()
    seam = compute_seam(mesh, rawf.singular)
    progress.report("seam done")
    cutm = compute_cut_mesh(mesh, seam)
    progress.report("cut mesh built")
    cfield, cmatch = compute_combed_field(mesh, rawf.field, rawf.matching, seam, N)
    ext = compute_extrinsic_field(mesh, cfield)
    progress.report("combed and extrinsic fields done")
    t["combinatorics"] = time.perf_counter() - t0
    rp = RosyParameterization(mesh, cutm, ext, rawf.singular, cmatch, seam, N, ratio)
    rp.seamless = seamless
    rp.localInjectivity = local_injectivity
    rp.roundSeams = round_seams
    t0 = time.perf_counter()
    rp.setup()
    progress.report("parameterization setup done")
    ok = rp.integ(literal=literal)
    t["igm"] = time.perf_counter() - t0
    uv = np.zeros((mesh.nF * 3, 2))
    for j in range(3):
        uv[j::3, 0] = rp.cfn[:, N * j]
        uv[j::3, 1] = rp.cfn[:, N * j + 1]
    ir = rp.ir
    return {
        "mesh": mesh, "cut": cutm, "field": rawf, "seam": seam, "combed": cfield,
        "combed_matching": cmatch, "ext": ext, "rp": rp, "success": ok, "uv": uv,
        "compressed": rawf.compressed, "singular": rawf.singular, "matching": rawf.matching,
        "nS": rp.nS, "nT": rp.nT, "fullx": rp.fullx, "x0": ir.x0,
        "decisions": getattr(ir, "decisions", []), "times": t,
    }
