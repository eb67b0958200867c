"""Benchmark: field + IGM solve on the 1.31M-face bumpy icosphere, FP32 multigrid inside
FP64 PCG vs plain FP64, both to a 1e-10 relative residual.

Workload (one "step"): icosphere level 8 (655,362 v / 1,310,720 f), radial displacement
0.05 sin(6 theta) sin(5 phi) plus N(0, sigma) radial jitter (seed 5) with sigma = 1e-3 * 2^(4-8)
= 6.25e-5: the level-4 jitter of 1e-3 scaled with the edge length, so every level carries the
same relative noise (at the unscaled 1e-3, a quarter of an L8 edge, the surface folds and the
smoothest field has 17,229 singularities: the "as_specified" entry reports that mesh's field
stage and the size of its IGM problem).  The example_igm pipeline end to
end: Smoothest 4-RoSy field (50 inverse-iteration solves), Curl matching, seam, cut, combing,
seamless IGM with iterative rounding (ratio 0.03).

value = field + IGM solve ms per step (median of the timed steps): everything in
mk_field_smallest_eig (multigrid setup + 50 PCG solves + normalisations) plus everything in
mk_iterative_rounding (kernel basis, operator assembly, multigrid setup, KKT, LM passes,
rounding loop).  ms_per_step is the whole pipeline (host mesh/seam/cut combinatorics too).
Inputs enter through host memory (the C ABI takes host arrays).

roofline: the kernel with the largest total time in one extra, event-timed step; achieved =
its algorithmic bytes summed over its launches / their summed time (hipEvents on the library
stream); traffic = HBM bytes per launch from the committed rocprofv3 --pmc summary when present.
field_spmv_fp64: the field operator's SpMV warm, after a cache flush, and streamed round robin
over copies whose total exceeds the Infinity Cache.

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
LEVEL = 8


def jitter_for(level):
    return 1e-3 * 2.0 ** (4 - level)


def workload(level=LEVEL, jitter=None):
    from metriko_amd.meshgen import bumpy_icosphere
    V, F = bumpy_icosphere(level, amp=0.05, jitter=jitter_for(level) if jitter is None else jitter, seed=5)
    g = np.random.Generator(np.random.PCG64(5))
    u0 = g.uniform(-1, 1, F.shape[0]) + 1j * g.uniform(-1, 1, F.shape[0])
    return V, F, u0


def solve_ms(info):
    return info["field"]["ms"] + info["field"]["setup_ms"] + info["igm"]["solve"]["ms"]


def cpu_baseline():
    """The CPU oracle (numpy/scipy SuperLU, one core) on a bounded sample: the same pipeline on
    the level-6 bumpy icosphere (81,920 faces, same relative jitter), field + IGM solve stages
    timed, scaled to 1,310,720 faces with the n^1.5 cost of sparse direct factorisation on a
    2-manifold mesh."""
    from oracle.pipeline import run_pipeline
    V, F, u0 = workload(6)
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


def cpu_baseline_full(level=LEVEL):
    """The CPU oracle (numpy/scipy SuperLU) once over the whole benchmark workload: the same
    mesh, field, seam, cut and seamless IGM as one GPU step, on the host this runs on, one
    thread (SuperLU has no threaded path; BLAS limited to one thread).  Prints each oracle stage
    as it ends (minutes in total) and returns the summary written to gpurun_out/."""
    import platform
    import scipy
    from threadpoolctl import threadpool_limits
    from oracle import progress
    from oracle.pipeline import run_pipeline
    progress.hook = lambda m: print(m, file=sys.stderr, flush=True)
    V, F, u0 = workload(level)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        r = run_pipeline(V, F, ratio=0.03, u0=u0)
        wall = time.perf_counter() - t0
    t = r["times"]
    return {
        "workload": f"bumpy icosphere L{level} ({V.shape[0]:,} v / {F.shape[0]:,} f), jitter N(0, {jitter_for(level):g}), "
                    f"field + seamless IGM, ratio 0.03 (the GPU benchmark's step)",
        "field_plus_igm_ms": round(1e3 * (t["field"] + t["igm"]), 1), "field_ms": round(1e3 * t["field"], 1),
        "igm_ms": round(1e3 * t["igm"], 1), "hmesh_ms": round(1e3 * t["hmesh"], 1),
        "combinatorics_ms": round(1e3 * t["combinatorics"], 1), "pipeline_wall_ms": round(1e3 * wall, 1),
        "nS": int(r["nS"]), "nT": int(r["nT"]), "success": bool(r["success"]),
        "rounding_decisions": len(r["decisions"]), "threads": 1, "nproc": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model, "python": platform.python_version(),
        "numpy": np.__version__, "scipy": scipy.__version__, "kind": "port",
    }


def traffic_for(key):
    """HBM bytes per launch of the kernel behind `key`, from profiles/traffic.json (written from
    the rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs), if present."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get(key)
    except (OSError, ValueError):
        return None


def as_specified(ctx, mk):
    """Configuration 5 as stated: the level-8 mesh with the unscaled N(0, 1e-3) jitter.  Its
    field stage is timed in both precisions (operator assembly on the device, multigrid set-up,
    50 solves), then the size of the IGM problem it leads to (host combinatorics); that IGM is
    not run (16.9k singularities, 87.6k constraint rows, ~77.6k rounding rounds)."""
    V, F, u0 = workload(LEVEL, jitter=1e-3)
    he = mk.hmesh_halfedges(V, F)
    res = {"jitter": 1e-3}
    x = None
    for name, prec in (("mixed", mk.PREC_MIXED_MG), ("fp64", mk.PREC_FP64)):
        t0 = time.perf_counter()
        L = ctx.field_laplacian(he["face_halfs"], he["twin"], he["half_face"], he["farg"], 4)
        xs, st = ctx.smallest_eig(L, he["face_area"], u0, 50, mk.opts(precision=prec))
        ms = 1e3 * (time.perf_counter() - t0)
        L.close()
        res[f"field_ms_{name}"] = round(ms, 1)
        res[f"field_iters_{name}"] = st["iterations"]
        res[f"field_rel_residual_{name}"] = st["rel_residual"]
        if x is None:
            x = xs
    c = mk.combinatorics(V, F, x)
    res.update({"singularities": c["nS"], "transitions": c["nT"], "constraint_rows": c["constraint_rows"],
                "seam_edges": int(c["seam"].sum()),
                "igm": "not run: the dense FullPivLU kernel of the constraint block and one rounding round per "
                       "integer variable (reference algorithm) do not finish at this size"})
    return res


def cold_spmv(ctx, mk, V, F, reps=24, copies=6):
    """The field operator's FP64 SpMV (complex, 1.31M rows, 5.24M entries, 147 MB algorithmic)
    timed with hipEvents in three ways:
    warm   back to back on one operator (the operator stays in the 256 MB Infinity Cache);
    flush  a 1 GiB buffer written (hipMemset) between repetitions: neither the Infinity Cache
           nor the L2s hold the operator, but the flush leaves them full of dirty lines that
           are written back while the SpMV runs;
    stream round robin over `copies` separate copies of the operator (6 x 125 MB of matrix,
           three times the Infinity Cache), so every SpMV reads its matrix from HBM, with no
           flush traffic: the operator-streaming rate of a solver whose working set exceeds
           the caches.
    frac_cold is the streaming rate over the 8 TB/s peak."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    flush_bytes = 1 << 30
    flush = C.c_void_p()
    if hip.hipMalloc(C.byref(flush), flush_bytes) != 0:
        raise RuntimeError("hipMalloc of the 1 GiB flush buffer failed")
    he = mk.hmesh_halfedges(V, F)
    Ls = [ctx.field_laplacian(he["face_halfs"], he["twin"], he["half_face"], he["farg"], 4) for _ in range(copies)]
    g = np.random.Generator(np.random.PCG64(7))
    x = g.standard_normal(F.shape[0]) + 1j * g.standard_normal(F.shape[0])
    out = {}
    for name in ("warm", "flush", "stream"):
        for L in Ls:
            L.apply(x)
        ctx.set_timing(True)
        ctx.timing_report(reset=True)
        for r in range(reps):
            if name == "flush":
                hip.hipMemset(flush, r & 0xff, flush_bytes)
                hip.hipDeviceSynchronize()
            (Ls[r % copies] if name == "stream" else Ls[0]).apply(x)
        rep = [k for k in ctx.timing_report(reset=True) if k["key"] == "field_L:spmv:fp64"]
        ctx.set_timing(False)
        k = rep[0]
        out[name] = {"us": round(1e3 * k["ms"] / k["calls"], 2), "bytes": round(k["bytes"] / k["calls"], 1),
                     "GBps": round(k["bytes"] / (k["ms"] * 1e-3) / 1e9, 1)}
    for L in Ls:
        L.close()
    hip.hipFree(flush)
    out["frac_cold"] = round(out["stream"]["GBps"] / PEAK_HBM_GBS, 4)
    return out


def measured_cpu_config():
    """The oracle's full run at the 1.31M-face workload on the GPU box's host, if its summary
    (profiles/cpu_baseline_L8.json, the summary `bench.py --cpu-baseline-full` writes) is present."""
    p = os.path.join(ROOT, "profiles", "cpu_baseline_L8.json")
    try:
        return json.load(open(p))
    except (OSError, ValueError):
        return None


def run(ctx, V, F, u0, o):
    t0 = time.perf_counter()
    out = ctx.pipeline(V, F, ratio=0.03, u0=u0, field_opts=o, igm_opts=o)
    return out, 1e3 * (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-as-specified", action="store_true")
    ap.add_argument("--no-cold-spmv", action="store_true")
    ap.add_argument("--level", type=int, default=LEVEL)
    ap.add_argument("--cpu-baseline-full", action="store_true",
                    help="time the CPU oracle once over the whole workload (minutes; no GPU use) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy a 1 GiB buffer three times (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported (one process drives one MI355X); got --gpus {args.gpus}. "
              f"The path shards by mesh with no communication; see DESIGN.md.", file=sys.stderr)
        sys.exit(2)
    if args.cpu_baseline_full:
        res = cpu_baseline_full(args.level)
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        with open(os.path.join(ROOT, "gpurun_out", f"cpu_baseline_L{args.level}.json"), "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps(res), flush=True)
        return
    import metriko_amd as mk
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    V, F, u0 = workload(args.level)
    mixed = mk.opts(precision=mk.PREC_MIXED_MG)
    fp64 = mk.opts(precision=mk.PREC_FP64)
    ctx = mk.Context(0)
    for w in range(args.warmup):
        run(ctx, V, F, u0, mixed)
        print(f"warmup {w} done", file=sys.stderr, flush=True)
    walls, solves, outs = [], [], []
    for s in range(args.steps):
        out, wall = run(ctx, V, F, u0, mixed)
        walls.append(wall)
        solves.append(solve_ms(out["info"]))
        outs.append(out)
        print(f"step {s}: wall {wall:.0f} ms solve {solves[-1]:.0f} ms", file=sys.stderr, flush=True)
    info = outs[-1]["info"]
    result = {
        "metric": "field+IGM solve ms & SpMV HBM GB/s @1.31M faces, FP32+refine vs FP64 @1e-10",
        "value": round(statistics.median(solves), 1), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(statistics.median(walls), 1), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64 (fp32 multigrid)", "data": "synthetic",
        "config": {"workload": f"bumpy icosphere L{args.level} ({V.shape[0]:,} v / {F.shape[0]:,} f), field + "
                               f"seamless IGM, ratio 0.03; DEVIATION from configuration 5: radial jitter "
                               f"N(0, {jitter_for(args.level):g}) (the level-4 jitter scaled with the edge length) "
                               f"instead of N(0, 1e-3), whose IGM is not run (see as_specified)",
                   "precision": "MK_PREC_MIXED_MG: FP64 flexible PCG, FP32 multigrid V-cycle, tol 1e-10",
                   "nS": info["nS"], "nT": info["nT"], "k": info["k"], "success": info["success"],
                   "field_ms": round(info["field"]["ms"] + info["field"]["setup_ms"], 1),
                   "field_iters": info["field"]["iterations"],
                   "igm_ms": round(info["igm"]["solve"]["ms"], 1), "igm_solves": info["igm"]["solve"]["solves"],
                   "igm_iters": info["igm"]["solve"]["iterations"],
                   "host_stages_ms": round(info["ms_hmesh"] + info["ms_matching"] + info["ms_seam"] +
                                           info["ms_cut"] + info["ms_setup"], 1)},
    }
    if not args.no_roofline:
        ctx.set_timing(True)
        ctx.timing_report(reset=True)
        run(ctx, V, F, u0, mixed)
        rep = ctx.timing_report(reset=True)
        ctx.set_timing(False)
        # per kernel key: algorithmic bytes summed over its launches (launches of one key can
        # move different amounts, e.g. the first Chebyshev step of a cycle) over the summed
        # hipEvent time of those launches
        dom = max(rep, key=lambda r: r["ms"])
        avg_ms = dom["ms"] / max(dom["calls"], 1)
        per_launch = dom["bytes"] / max(dom["calls"], 1)
        achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
        result["roofline"] = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_for(dom["key"]),
            "kernel": dom["key"], "avg_us": round(avg_ms * 1e3, 2), "bytes_per_launch": round(per_launch, 1),
            "launches": dom["calls"], "rows": dom["rows"]}
        top = sorted(rep, key=lambda r: -r["ms"])[:8]
        result["top_kernels"] = [{"key": r["key"], "ms": round(r["ms"], 1), "calls": r["calls"],
                                  "GBps": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 1)} for r in top]
    if not args.no_fp64:
        o64, w64 = run(ctx, V, F, u0, fp64)
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
                # the same (index, value) decisions; the order in which candidates whose
                # distances to the grid agree to round-off are fixed may differ
                "decisions_equal": sorted(a["decisions"]) == sorted(b["decisions"]),
                "decisions_same_order": a["decisions"] == b["decisions"]},
        }
        result["speedup_mixed_vs_fp64"] = round(result["fp64"]["solve_ms"] / result["value"], 3)
    result["breakdown_ms"] = {
        "field_setup": round(info["field"]["setup_ms"], 1), "field_solves": round(info["field"]["ms"], 1),
        "igm_host_setup": round(info["igm"]["setup_ms"], 1), "igm_kkt": round(info["igm"]["kkt_ms"], 1),
        "igm_naive": round(info["igm"]["naive_ms"], 1), "igm_rounding": round(info["igm"]["rounding_ms"], 1)}
    if not args.no_as_specified and args.level == LEVEL:
        result["as_specified"] = as_specified(ctx, mk)
    if not args.no_cold_spmv:
        result["field_spmv_fp64"] = cold_spmv(ctx, mk, V, F)
    ctx.close()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
        full = measured_cpu_config()
        if full:
            result["cpu_baseline"]["measured_full_config"] = full
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
