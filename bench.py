"""Benchmark: field + IGM solve on the 1.31M-face bumpy icosphere, FP32+FP64-refinement vs FP64.

Workload (one "step"): icosphere level 8 (655,362 v / 1,310,720 f), radial displacement
0.05 sin(6 theta) sin(5 phi) plus N(0, 1e-3) radial jitter (seed 5); the example_igm
pipeline end to end: Smoothest 4-RoSy field (50 inverse-iteration solves), Curl matching,
seam, cut, combing, seamless IGM with iterative rounding (ratio 0.03).

value = field + IGM solve ms per step (median over the timed steps), FP32 inner solves with
FP64 refinement to a 1e-10 FP64 relative residual.  "Solve" = everything in
mk_field_smallest_eig (multigrid setup + 50 PCG solves + normalisations) plus everything in
mk_iterative_rounding (kernel basis, operator assembly, multigrid setup, KKT, LM passes,
rounding loop); the host-side mesh/seam/cut combinatorics are reported separately in
ms_per_step (whole pipeline).  Inputs enter through host memory (the C ABI takes host
arrays), so the timed region includes their transfer.

Prints ONE JSON line.  Only --gpus 1 is supported (one process, one GPU).
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def workload():
    from metriko_amd.meshgen import bumpy_icosphere
    V, F = bumpy_icosphere(8, amp=0.05, jitter=1e-3, seed=5)
    g = np.random.Generator(np.random.PCG64(5))
    u0 = g.uniform(-1, 1, F.shape[0]) + 1j * g.uniform(-1, 1, F.shape[0])
    return V, F, u0


def solve_ms(info):
    return info["field"]["ms"] + info["field"]["setup_ms"] + info["igm"]["solve"]["ms"]


def cpu_baseline():
    """The CPU oracle (numpy/scipy SuperLU, one core) on a bounded sample: the same
    pipeline on the level-6 bumpy icosphere (81,920 faces), field + IGM solve stages timed,
    scaled to 1,310,720 faces with the n^1.5 cost of sparse direct factorisation on a
    2-manifold mesh."""
    from metriko_amd.meshgen import bumpy_icosphere
    from oracle.pipeline import run_pipeline
    V, F = bumpy_icosphere(6, amp=0.05, jitter=1e-3, seed=5)
    g = np.random.Generator(np.random.PCG64(5))
    u0 = g.uniform(-1, 1, F.shape[0]) + 1j * g.uniform(-1, 1, F.shape[0])
    t0 = time.perf_counter()
    r = run_pipeline(V, F, ratio=0.03, u0=u0)
    wall = time.perf_counter() - t0
    sample_ms = 1e3 * (r["times"]["field"] + r["times"]["igm"])
    scale = (1310720 / F.shape[0]) ** 1.5
    return {
        "value": round(sample_ms * scale, 1), "unit": "ms", "cores": 1, "kind": "port",
        "sample": (f"oracle field+IGM solve stages on the level-6 bumpy icosphere (81,920 faces): "
                   f"{sample_ms:.0f} ms measured (pipeline wall {wall:.1f} s), scaled x{scale:.0f} "
                   f"= (1,310,720/81,920)^1.5 to the 1.31M-face workload"),
    }


def traffic_for(kernel_symbol):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(kernel_symbol)
    except Exception:
        return None


# kernel symbol each timing key maps to (rocprofv3 names)
SYMBOL = {
    ("cg_spmv", True, "fp32"): "cg_spmv<float, true>",
    ("cg_spmv", False, "fp32"): "cg_spmv<float, false>",
    ("cg_spmv", True, "fp64"): "cg_spmv<double, true>",
    ("cg_spmv", False, "fp64"): "cg_spmv<double, false>",
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported (one process drives one MI355X); got --gpus {args.gpus}. "
              f"The batched and multi-mesh paths shard by mesh with no communication; see DESIGN.md.",
              file=sys.stderr)
        sys.exit(2)
    import metriko_amd as mk
    V, F, u0 = workload()
    mixed = mk.opts(precision=mk.PREC_MIXED)
    fp64 = mk.opts(precision=mk.PREC_FP64)
    ctx = mk.Context(0)
    for _ in range(args.warmup):
        ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=mixed, igm_opts=mixed)
        print(f"warmup done", file=sys.stderr, flush=True)
    ctx.set_timing(True)
    ctx.timing_report(reset=True)
    walls, solves, infos, outs = [], [], [], []
    for s in range(args.steps):
        t0 = time.perf_counter()
        out = ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=mixed, igm_opts=mixed)
        walls.append(1e3 * (time.perf_counter() - t0))
        solves.append(solve_ms(out["info"]))
        infos.append(out["info"])
        outs.append(out)
        print(f"step {s}: wall {walls[-1]:.0f} ms solve {solves[-1]:.0f} ms", file=sys.stderr, flush=True)
    rep = ctx.timing_report(reset=True)
    ctx.set_timing(False)
    # dominant SpMV kernel by total event time
    dom = max(rep, key=lambda r: r["ms"])
    op, kern, prec = dom["key"].split(":")
    avg_ms = dom["ms"] / max(dom["calls"], 1)
    achieved = dom["bytes_per_call"] / (avg_ms * 1e-3) / 1e9
    cplx = op.startswith("field")
    sym = SYMBOL.get((kern, cplx, prec), kern)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_for(sym),
                "kernel": sym, "key": dom["key"], "avg_us": round(avg_ms * 1e3, 2),
                "bytes_per_launch": dom["bytes_per_call"], "launches": dom["calls"]}
    info = infos[-1]
    result = {
        "metric": "field+IGM solve ms & SpMV HBM GB/s @1.31M faces, FP32+refine vs FP64 @1e-10",
        "value": round(statistics.median(solves), 1), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(statistics.median(walls), 1), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64-refine", "data": "synthetic",
        "config": {"workload": "bumpy icosphere L8 (655,362 v / 1,310,720 f), field + seamless IGM, ratio 0.03",
                   "nS": info["nS"], "nT": info["nT"], "k": info["k"],
                   "field_ms": round(info["field"]["ms"] + info["field"]["setup_ms"], 1),
                   "field_iters": info["field"]["iterations"],
                   "igm_ms": round(info["igm"]["solve"]["ms"], 1), "igm_solves": info["igm"]["solve"]["solves"],
                   "igm_iters": info["igm"]["solve"]["iterations"],
                   "rel_residual": info["igm"]["solve"]["rel_residual"]},
        "roofline": roofline,
    }
    if not args.no_fp64:
        t0 = time.perf_counter()
        o64 = ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=fp64, igm_opts=fp64)
        w64 = 1e3 * (time.perf_counter() - t0)
        a, b = outs[-1], o64
        c = np.vdot(b["compressed"], a["compressed"])
        c = c / abs(c)
        field_err = np.linalg.norm(a["compressed"] - c * b["compressed"]) / np.linalg.norm(b["compressed"])
        result["fp64"] = {
            "solve_ms": round(solve_ms(o64["info"]), 1), "ms_per_step": round(w64, 1),
            "field_iters": o64["info"]["field"]["iterations"], "igm_iters": o64["info"]["igm"]["solve"]["iterations"],
            "mixed_vs_fp64": {
                "field_phase_err": float(field_err),
                "uv_rel_err": float(np.abs(a["uv"] - b["uv"]).max() / np.abs(b["uv"]).max()),
                "singular_equal": bool(np.array_equal(a["singular"], b["singular"])),
                "seam_equal": bool(np.array_equal(a["seam"], b["seam"])),
                "decisions_equal": a["decisions"] == b["decisions"]},
        }
        result["speedup_mixed_vs_fp64"] = round(result["fp64"]["solve_ms"] / result["value"], 3)
    ctx.close()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
