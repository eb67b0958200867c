"""Writes the oracle's outputs on the reference's own input fixture as regression vectors.

The reference ships no expected outputs; these fixtures freeze the oracle (pinned by
tests/test_oracle_pinning.py) on example/models/icosphere.obj with the example_igm flag
set and ratio 0.03, so later changes to the oracle or the GPU path are caught exactly.
Run: python tests/golden/make_golden.py            (icosphere.obj fixture)
     python tests/golden/make_golden.py --torus    (204,800-face torus, about two minutes)
     python tests/golden/make_golden.py --l8       (the 1.31M-face benchmark mesh, about 17 minutes)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from metriko_amd.meshgen import read_obj  # noqa: E402
from oracle.field import eigen_random_complex  # noqa: E402
from oracle.pipeline import run_pipeline  # noqa: E402
from tests.helpers import cut_hash, torus_config  # noqa: E402


def make_compact(name, V, F, u0, face_step, corner_step):
    """The oracle's pipeline on (V, F, u0), kept compactly: the field on every face_step-th face,
    uv on every corner_step-th corner, the singular vertices, the seam edges, a checksum of the
    cut-mesh indices and the rounding decisions."""
    r = run_pipeline(V, F, ratio=0.03, u0=u0)
    dec = np.array([(i, v) for (i, v, _) in r["decisions"]], dtype=np.int64)
    margin = np.array([m for (_, _, m) in r["decisions"]])
    sf = np.arange(0, F.shape[0], face_step)         # sampled faces
    sc = np.arange(0, 3 * F.shape[0], corner_step)   # sampled corners
    sing = np.nonzero(r["singular"])[0]
    np.savez_compressed(
        os.path.join(HERE, name),
        field_faces=sf, field=r["compressed"][sf], singular_v=sing.astype(np.int32),
        singular_idx=r["singular"][sing].astype(np.int32), seam_edges=np.nonzero(r["seam"])[0].astype(np.int32),
        cut_hash=np.int64(cut_hash(r["cut"].idx)), cutV=np.int64(r["cut"].nV), uv_corners=sc,
        uv=r["uv"].reshape(-1, 2)[sc], decisions=dec, decision_margin=margin, nS=r["nS"], nT=r["nT"],
        success=np.int64(bool(r["success"])))
    print("wrote", name, r["nS"], r["nT"], len(dec))


def make_torus():
    V, F, u0 = torus_config()
    make_compact("torus_256x400_oracle.npz", V, F, u0, 50, 150)


def make_l8():
    from bench import workload
    V, F, u0 = workload(8)
    make_compact("bumpy_L8_oracle.npz", V, F, u0, 500, 1500)


def main():
    if "--torus" in sys.argv:
        make_torus()
        return
    if "--l8" in sys.argv:
        make_l8()
        return
    V, F = read_obj(os.path.join(HERE, "icosphere.obj"))
    u0 = eigen_random_complex(F.shape[0])
    r = run_pipeline(V, F, ratio=0.03, u0=u0)
    dec = np.array([(i, v) for (i, v, _) in r["decisions"]], dtype=np.int64)
    margin = np.array([m for (_, _, m) in r["decisions"]])
    np.savez_compressed(
        os.path.join(HERE, "icosphere_obj_oracle.npz"),
        compressed=r["compressed"], singular=r["singular"].astype(np.int32),
        seam_edges=np.nonzero(r["seam"])[0].astype(np.int32), cut_idx=r["cut"].idx.astype(np.int32),
        combed_matching=r["combed_matching"].astype(np.int32), decisions=dec, decision_margin=margin,
        uv=r["uv"], nS=r["nS"], nT=r["nT"])
    print("wrote icosphere_obj_oracle.npz", r["nS"], r["nT"], len(dec))


if __name__ == "__main__":
    main()
