"""metriko_amd: MI355X (gfx950) HIP implementation of metriko's field + IGM hot path.

Thin ctypes binding over ``libmetriko_hip.so`` (C ABI declared in
``include/metriko_hip.h``).  The binding is plumbing for tests and the benchmark:
all numerics run in the library.  Importing it never falls back to anything else:
if the library is missing, ``lib()`` raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmetriko_hip.so")
_lib = None

PREC_FP64, PREC_MIXED, PREC_MIXED_MG = 0, 1, 2
FMT_CSR, FMT_SELL = 0, 1
PRECOND_NONE, PRECOND_JACOBI, PRECOND_AMG = 0, 1, 2

EXPORTED = [
    "mk_status_str", "mk_default_opts", "mk_init", "mk_finalize", "mk_last_error", "mk_set_timing", "mk_set_graphs", "mk_timing_report",
    "mk_op_create_real", "mk_op_create_complex", "mk_op_destroy", "mk_op_apply", "mk_spd_solve",
    "mk_field_laplacian", "mk_op_export_csr", "mk_field_smallest_eig", "mk_iterative_rounding",
    "mk_pipeline_run", "mk_combinatorics", "mk_hmesh_export", "mk_hmesh_halfedges", "mk_mg_host_pcg",
    "mk_spd_solve_many", "mk_field_solve", "mk_curvature_dir", "mk_field_matching", "mk_autotune_spmv",
    "mk_batch_run", "mk_igm_injectivity_check", "mk_qex_extract", "mk_motorcycle_graph", "mk_tmesh_build", "mk_qgp_quantize", "mk_read_obj", "mk_cfn_write_json",
    "mk_cfn_read_json", "mk_keep_igm_state", "mk_igm_state_take", "mk_igm_state_destroy", "mk_igm_state_sizes",
    "mk_igm_sized_uv", "mk_igm_sized_uv_grad", "mk_spd_solve_noalloc", "mk_set_naive_iter_callback",
]
# iterative_rounding's iter_cb across the C ABI: (user, fullx, nfull) -> nonzero to stop
NAIVE_ITER_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int)
FIELD_SMOOTHEST, FIELD_CURVATURE_ALIGNED = 0, 1
MATCH_PRINCIPAL, MATCH_CURL = 0, 1


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("tol", C.c_double), ("maxiter", C.c_int), ("format", C.c_int),
                ("slice_height", C.c_int), ("sigma", C.c_int), ("block_size", C.c_int), ("precond", C.c_int),
                ("inner_tol", C.c_double), ("max_refine", C.c_int), ("check_every", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int), ("solves", C.c_int),
                ("rel_residual", C.c_double), ("ms", C.c_double), ("spmv_ms", C.c_double),
                ("spmv_calls", C.c_int64), ("spmv_bytes", C.c_double), ("setup_ms", C.c_double),
                ("mg_levels", C.c_int)]


class IgmFlags(C.Structure):
    _fields_ = [("seamless", C.c_int), ("round_seams", C.c_int), ("local_injectivity", C.c_int),
                ("verbose", C.c_int)]


class IgmStats(C.Structure):
    _fields_ = [("rounds", C.c_int), ("lm_rounds", C.c_int), ("k", C.c_int), ("kernel_rank", C.c_int),
                ("n_fixed", C.c_int), ("null_vectors", C.c_int), ("batched_solves", C.c_int),
                ("batch_iterations", C.c_int), ("device_order", C.c_int), ("kkt_rel_residual", C.c_double),
                ("setup_ms", C.c_double),
                ("kkt_ms", C.c_double), ("naive_ms", C.c_double), ("rounding_ms", C.c_double), ("solve", Stats)]


class QexCounts(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n_vq", "n_eq", "n_fq", "n_ports", "n_edges", "n_faces")]


class PipelineInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nV", "nF", "nE", "cutV", "nS", "nT", "k", "nfullx", "ndec", "success")] + \
               [(n, C.c_double) for n in ("ms_hmesh", "ms_field_asm", "ms_field_solve", "ms_matching", "ms_seam",
                                          "ms_cut", "ms_setup", "ms_integ", "ms_total")] + \
               [("field", Stats), ("igm", IgmStats)]


class BatchStats(C.Structure):
    _fields_ = [("meshes", C.c_int), ("faces", C.c_int64), ("verts", C.c_int64),
                ("field_iterations", C.c_int), ("poisson_iterations", C.c_int),
                ("field_rel_residual", C.c_double), ("poisson_rel_residual", C.c_double),
                ("ms_host_setup", C.c_double), ("ms_field", C.c_double), ("ms_poisson", C.c_double),
                ("ms_total", C.c_double), ("tuned_slice_height", C.c_int), ("tuned_sigma", C.c_int),
                ("tuned_block", C.c_int), ("tuned_spmv_us", C.c_double),
                ("tuned_poisson_slice_height", C.c_int), ("tuned_poisson_sigma", C.c_int),
                ("tuned_poisson_block", C.c_int), ("tuned_poisson_spmv_us", C.c_double),
                ("graph_replays", C.c_int64), ("graph_captures", C.c_int64), ("eager_batches", C.c_int64)]


def struct_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = struct_dict(v) if isinstance(v, C.Structure) else v
    return out


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)


def lib():
    """Load libmetriko_hip.so (built by ``metriko_amd/build.py``); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"metriko_amd: {LIB_PATH} is not built (run python -m metriko_amd.build)")
    L = C.CDLL(LIB_PATH)
    L.mk_status_str.restype = C.c_char_p
    L.mk_status_str.argtypes = [C.c_int]
    L.mk_default_opts.argtypes = [C.POINTER(SolveOpts)]
    L.mk_init.argtypes = [C.POINTER(_P), C.c_int]
    L.mk_finalize.argtypes = [_P]
    L.mk_last_error.argtypes = [_P, C.c_char_p, C.c_int]
    L.mk_set_timing.argtypes = [_P, C.c_int]
    L.mk_set_graphs.argtypes = [_P, C.c_int]
    L.mk_timing_report.argtypes = [_P, C.c_int, C.c_char_p, C.c_int]
    L.mk_op_create_real.argtypes = [_P, C.c_int, _I64, _I32, _D, C.POINTER(SolveOpts), C.POINTER(_P)]
    L.mk_op_create_complex.argtypes = [_P, C.c_int, _I64, _I32, _D, C.POINTER(SolveOpts), C.POINTER(_P)]
    L.mk_op_destroy.argtypes = [_P]
    L.mk_op_apply.argtypes = [_P, _P, _D, _D]
    L.mk_spd_solve.argtypes = [_P, _P, _D, _D, C.POINTER(SolveOpts), C.POINTER(Stats)]
    L.mk_spd_solve_noalloc.argtypes = [_P, _P, C.c_void_p, C.c_void_p, C.POINTER(SolveOpts), C.POINTER(Stats)]
    L.mk_spd_solve_many.argtypes = [_P, _P, C.c_int, _D, _D, C.POINTER(SolveOpts), C.POINTER(Stats)]
    L.mk_field_laplacian.argtypes = [_P, C.c_int, C.c_int, _I32, _I32, _I32, _D, C.c_int, C.POINTER(SolveOpts),
                                     C.POINTER(_P)]
    L.mk_op_export_csr.argtypes = [_P, _P, _I64, _I64, _I32, _D]
    L.mk_field_smallest_eig.argtypes = [_P, _P, _D, _D, C.c_int, C.POINTER(SolveOpts), _D, C.POINTER(Stats)]
    L.mk_iterative_rounding.argtypes = [
        _P, C.c_int, _I32, _D, C.c_int, _I32, C.c_int, _I32, C.c_double,
        C.c_int, C.c_int, _I64, _I32, _D, C.c_int, C.c_int, _I64, _I32, _D,
        C.c_int, C.c_int, C.c_int, C.POINTER(IgmFlags), _D, C.POINTER(SolveOpts), _D, C.POINTER(C.c_int),
        C.c_int, _I32, _I32, _D, C.POINTER(C.c_int), C.POINTER(IgmStats)]
    L.mk_pipeline_run.argtypes = [_P, C.c_int, C.c_int, _D, _I32, C.c_int, C.c_double, _D, C.c_int,
                                  C.POINTER(SolveOpts), C.POINTER(SolveOpts), _D, _I32, _I32, _U8, _I32, _D,
                                  C.c_int, _D, C.c_int, _I32, _I32, C.POINTER(PipelineInfo)]
    L.mk_combinatorics.argtypes = [C.c_int, C.c_int, _D, _I32, C.c_int, _D, _I32, _I32, _I32, _U8, _I32, _I32, _D]
    L.mk_hmesh_export.argtypes = [C.c_int, C.c_int, _D, _I32, C.POINTER(C.c_int), _I32, _I32, _I32, _I32, _D, _D]
    L.mk_hmesh_halfedges.argtypes = [C.c_int, C.c_int, _D, _I32, _I32, _I32, _I32, _D, _D]
    L.mk_field_solve.argtypes = [_P, C.c_int, C.c_int, _D, _I32, C.c_int, C.c_int, _D, C.c_int,
                                 C.POINTER(SolveOpts), _D, C.POINTER(Stats)]
    L.mk_curvature_dir.argtypes = [C.c_int, C.c_int, _D, _I32, _D]
    L.mk_field_matching.argtypes = [C.c_int, C.c_int, _D, _I32, C.c_int, _D, C.c_int, _I32, _I32, _D]
    L.mk_autotune_spmv.argtypes = [_P, C.c_int, _I64, _I32, _D, C.c_int, C.POINTER(SolveOpts), C.POINTER(SolveOpts),
                                   C.POINTER(C.c_double)]
    L.mk_batch_run.argtypes = [_P, C.c_int, _I32, _I32, C.POINTER(_D), C.POINTER(_I32), C.c_int, C.c_int,
                               C.POINTER(_D), C.POINTER(SolveOpts), C.POINTER(_D), C.POINTER(_D),
                               C.POINTER(BatchStats)]
    L.mk_igm_injectivity_check.argtypes = [_P, C.c_int, C.c_int, _I32, C.c_int, C.c_int, _I64, _I32, _D]
    L.mk_keep_igm_state.argtypes = [_P, C.c_int]
    L.mk_set_naive_iter_callback.argtypes = [_P, NAIVE_ITER_CB, C.c_void_p]
    L.mk_motorcycle_graph.argtypes = [_P, C.c_int, C.c_int, _D, _I32, _I32, _I32, _D, C.c_int, _I32, _D, C.c_int,
                                      _I32, _D, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mk_tmesh_build.argtypes = [_P, C.c_int, C.c_int, _D, _I32, _I32, _I32, _D, C.c_int, _I32, _I32, C.c_int, _I32,
                                 C.c_int, _I32, _I32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mk_qgp_quantize.argtypes = [_P, C.c_int, C.c_int, _D, _I32, _I32, _I32, _D, C.c_int, _D, _D,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mk_read_obj.argtypes = [C.c_char_p, C.c_int, C.c_int, _D, _I32, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mk_cfn_write_json.argtypes = [C.c_char_p, C.c_int, _D]
    L.mk_cfn_read_json.argtypes = [C.c_char_p, C.c_int, _D, C.POINTER(C.c_int)]
    L.mk_qex_extract.argtypes = [_P, C.c_int, C.c_int, _D, _I32, C.c_int, _I32, _I32, _D, _D, C.c_int, _D, _D, _I32,
                                 C.c_int, _I32, _D, C.c_int, _I32, C.c_int, _I32, C.POINTER(QexCounts)]
    L.mk_igm_state_take.argtypes = [_P, C.POINTER(_P)]
    L.mk_igm_state_destroy.argtypes = [_P]
    L.mk_igm_state_sizes.argtypes = [_P] + [C.POINTER(C.c_int)] * 5
    L.mk_igm_sized_uv.argtypes = [_P, _P, _D, _D, _D]
    L.mk_igm_sized_uv_grad.argtypes = [_P, _P, _D, _D, _D]
    L.mk_mg_host_pcg.argtypes = [C.c_int, _I64, _I32, _D, C.c_int, _D, C.c_int, _D, _D, C.c_double, C.c_int,
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    for name in EXPORTED:
        if name not in ("mk_status_str", "mk_default_opts"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class MetrikoError(RuntimeError):
    def __init__(self, status, what, detail=""):
        self.status = status
        super().__init__(f"{what}: {lib().mk_status_str(status).decode()} {detail}".strip())


def opts(**kw):
    """mk_solve_opts with the library defaults, overridden by keyword arguments."""
    o = SolveOpts()
    lib().mk_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _mesh_arrays(V, F):
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    assert V.ndim == 2 and V.shape[1] == 3 and F.ndim == 2 and F.shape[1] == 3
    return V, F


def hmesh_export(V, F):
    """Topology/geometry of the library's Hmesh (host code; no GPU needed)."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    nE = C.c_int(0)
    s = L.mk_hmesh_export(nV, nF, _ptr(V, _D), _ptr(F, _I32), C.byref(nE), None, None, None, None, None, None)
    if s:
        raise MetrikoError(s, "mk_hmesh_export")
    nE = nE.value
    ev = np.zeros((nE, 2), np.int32)
    ef = np.zeros((nE, 2), np.int32)
    fe = np.zeros((nF, 3), np.int32)
    vh = np.zeros(nV, np.int32)
    fa = np.zeros(2 * nE)
    ad = np.zeros(nV)
    n2 = C.c_int(0)
    s = L.mk_hmesh_export(nV, nF, _ptr(V, _D), _ptr(F, _I32), C.byref(n2), _ptr(ev, _I32), _ptr(ef, _I32),
                          _ptr(fe, _I32), _ptr(vh, _I32), _ptr(fa, _D), _ptr(ad, _D))
    if s:
        raise MetrikoError(s, "mk_hmesh_export")
    return {"nE": nE, "edge2vert": ev, "edge2face": ef, "face2edge": fe, "vert2half": vh, "farg": fa,
            "angle_defect": ad}


def hmesh_halfedges(V, F):
    """face_halfs, twin, half_face, farg, face_area of the library's Hmesh (host code)."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    nH = 2 * hmesh_export(V, F)["nE"]
    fh = np.zeros((nF, 3), np.int32)
    tw = np.zeros(nH, np.int32)
    hf = np.zeros(nH, np.int32)
    fa = np.zeros(nH)
    ar = np.zeros(nF)
    s = L.mk_hmesh_halfedges(nV, nF, _ptr(V, _D), _ptr(F, _I32), _ptr(fh, _I32), _ptr(tw, _I32), _ptr(hf, _I32),
                             _ptr(fa, _D), _ptr(ar, _D))
    if s:
        raise MetrikoError(s, "mk_hmesh_halfedges")
    return {"face_halfs": fh, "twin": tw, "half_face": hf, "farg": fa, "face_area": ar}


def mg_host_pcg(A, b, tol=1e-10, maxit=500, nullvec=None):
    """Host-only multigrid-preconditioned CG on a CSR matrix (same hierarchy as the GPU path).
    `nullvec` (n x k, real operators only) are near-null vectors for the aggregation.
    Returns (x, iterations, levels, relative residual)."""
    import scipy.sparse as sp
    L = lib()
    cplx = np.iscomplexobj(A) or np.iscomplexobj(b)
    A = sp.csr_matrix(A, dtype=np.complex128 if cplx else np.float64)
    A.sort_indices()
    ptr = A.indptr.astype(np.int64)
    col = A.indices.astype(np.int32)
    if cplx:
        val = np.ascontiguousarray(np.stack([A.data.real, A.data.imag], 1).reshape(-1))
        bb = np.ascontiguousarray(np.stack([np.real(b), np.imag(b)], 1).reshape(-1))
    else:
        val = A.data.astype(np.float64)
        bb = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(bb)
    nnull, nv = 0, None
    if nullvec is not None:
        z = np.asarray(nullvec).reshape(A.shape[0], -1)
        nnull = z.shape[1]
        if cplx:
            z = z.astype(np.complex128).T  # column-major, complex interleaved
            nv = np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(-1))
        else:
            nv = np.ascontiguousarray(z.astype(np.float64).T)
    it, lv, rr = C.c_int(0), C.c_int(0), C.c_double(0)
    s = L.mk_mg_host_pcg(A.shape[0], _ptr(ptr, _I64), _ptr(col, _I32), _ptr(val, _D), int(cplx),
                         _ptr(nv, _D) if nnull else None, nnull, _ptr(bb, _D), _ptr(x, _D), tol, maxit,
                         C.byref(it), C.byref(lv), C.byref(rr))
    if s:
        raise MetrikoError(s, "mk_mg_host_pcg")
    if cplx:
        x = x[0::2] + 1j * x[1::2]
    return x, it.value, lv.value, rr.value


def motorcycle_graph(V, F, matching, singular, uv):
    """QGP motorcycle graph (mk_motorcycle_graph, host code) on corner uv [3 nF, 2]: ports
    (id, vertex, face, prev, next) with their uv / direction, and every curve's segments
    (curve, face, from/to type, from/to crash curve) with their end points."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    mt = np.ascontiguousarray(matching, np.int32)
    sg = np.ascontiguousarray(singular, np.int32)
    u = np.ascontiguousarray(np.asarray(uv, np.float64).reshape(-1))
    npo, nse = C.c_int(), C.c_int()
    args = (None, nV, nF, _ptr(V, _D), _ptr(F, _I32), _ptr(mt, _I32), _ptr(sg, _I32), _ptr(u, _D))
    s = L.mk_motorcycle_graph(*args, 0, None, None, 0, None, None, C.byref(npo), C.byref(nse))
    if s:
        raise MetrikoError(s, "mk_motorcycle_graph")
    pid, puv = np.zeros((npo.value, 5), np.int32), np.zeros((npo.value, 4))
    sid, suv = np.zeros((nse.value, 6), np.int32), np.zeros((nse.value, 4))
    s = L.mk_motorcycle_graph(*args, npo.value, _ptr(pid, _I32), _ptr(puv, _D), nse.value, _ptr(sid, _I32),
                              _ptr(suv, _D), C.byref(npo), C.byref(nse))
    if s:
        raise MetrikoError(s, "mk_motorcycle_graph")
    return {"port_ids": pid, "port_uvdir": puv, "seg_ids": sid, "seg_uv": suv}


def tmesh_build(V, F, matching, singular, uv):
    """QGP T-mesh (mk_tmesh_build, host code): T-edges (curve, first, last segment),
    th2sing, and the T-quads as lists of (half-edge, side)."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    mt = np.ascontiguousarray(matching, np.int32)
    sg = np.ascontiguousarray(singular, np.int32)
    u = np.ascontiguousarray(np.asarray(uv, np.float64).reshape(-1))
    ne, nq, nt = C.c_int(), C.c_int(), C.c_int()
    args = (None, nV, nF, _ptr(V, _D), _ptr(F, _I32), _ptr(mt, _I32), _ptr(sg, _I32), _ptr(u, _D))
    s = L.mk_tmesh_build(*args, 0, None, None, 0, None, 0, None, None, C.byref(ne), C.byref(nq), C.byref(nt))
    if s:
        raise MetrikoError(s, "mk_tmesh_build")
    te, ts = np.zeros((ne.value, 3), np.int32), np.zeros(2 * ne.value, np.int32)
    qp, qt, qs = np.zeros(nq.value + 1, np.int32), np.zeros(nt.value, np.int32), np.zeros(nt.value, np.int32)
    s = L.mk_tmesh_build(*args, ne.value, _ptr(te, _I32), _ptr(ts, _I32), nq.value, _ptr(qp, _I32), nt.value,
                         _ptr(qt, _I32), _ptr(qs, _I32), C.byref(ne), C.byref(nq), C.byref(nt))
    if s:
        raise MetrikoError(s, "mk_tmesh_build")
    quads = [(qt[qp[q]:qp[q + 1]].tolist(), qs[qp[q]:qp[q + 1]].tolist()) for q in range(nq.value)]
    return {"tedges": te, "th2sing": ts, "tquads": quads}


def qgp_quantize(V, F, matching, singular, uv):
    """QGP quantization (mk_qgp_quantize, host code): ideal T-edge lengths R, integer lengths X
    and whether X passed the reference's checks."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    mt = np.ascontiguousarray(matching, np.int32)
    sg = np.ascontiguousarray(singular, np.int32)
    u = np.ascontiguousarray(np.asarray(uv, np.float64).reshape(-1))
    ne, ok = C.c_int(), C.c_int()
    args = (None, nV, nF, _ptr(V, _D), _ptr(F, _I32), _ptr(mt, _I32), _ptr(sg, _I32), _ptr(u, _D))
    s = L.mk_qgp_quantize(*args, 0, None, None, C.byref(ne), None)
    if s:
        raise MetrikoError(s, "mk_qgp_quantize")
    R, X = np.zeros(ne.value), np.zeros(ne.value)
    s = L.mk_qgp_quantize(*args, ne.value, _ptr(R, _D), _ptr(X, _D), C.byref(ne), C.byref(ok))
    if s:
        raise MetrikoError(s, "mk_qgp_quantize")
    return {"ideal": R, "lengths": X, "valid": bool(ok.value)}


def read_obj(path):
    """Triangle OBJ through the library's reader (mk_read_obj): V [nV, 3], F [nF, 3]."""
    L = lib()
    nV, nF = C.c_int(), C.c_int()
    p = str(path).encode()
    s = L.mk_read_obj(p, 0, 0, None, None, C.byref(nV), C.byref(nF))
    if s:
        raise MetrikoError(s, "mk_read_obj")
    V = np.zeros((nV.value, 3))
    F = np.zeros((nF.value, 3), np.int32)
    s = L.mk_read_obj(p, nV.value, nF.value, _ptr(V, _D), _ptr(F, _I32), C.byref(nV), C.byref(nF))
    if s:
        raise MetrikoError(s, "mk_read_obj")
    return V, F


def write_cfn_json(path, cfn):
    """Corner uv (complex [3 nF] or real [3 nF, 2]) to the {"cfn": [...]} dump (mk_cfn_write_json)."""
    c = np.asarray(cfn)
    a = np.ascontiguousarray(np.stack([c.real, c.imag], 1) if np.iscomplexobj(c) else c, np.float64).reshape(-1)
    s = lib().mk_cfn_write_json(str(path).encode(), a.size // 2, _ptr(a, _D))
    if s:
        raise MetrikoError(s, "mk_cfn_write_json")


def read_cfn_json(path):
    """The {"cfn": [...]} dump as complex corner uv (mk_cfn_read_json)."""
    L = lib()
    n = C.c_int()
    p = str(path).encode()
    s = L.mk_cfn_read_json(p, 0, None, C.byref(n))
    if s:
        raise MetrikoError(s, "mk_cfn_read_json")
    a = np.zeros(2 * n.value)
    s = L.mk_cfn_read_json(p, n.value, _ptr(a, _D), C.byref(n))
    if s:
        raise MetrikoError(s, "mk_cfn_read_json")
    return a[0::2] + 1j * a[1::2]


def qex_extract(V, F, matching, singular, uv, N=4):
    """QEx quad extraction (mk_qex_extract, host code; no GPU needed) on corner uv [3 nF, 2]:
    sanitized uv, q-vertices (kind 0/1/2 = on a vertex / edge / face), ports, edges, quads."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    mt = np.ascontiguousarray(matching, np.int32)
    sg = np.ascontiguousarray(singular, np.int32)
    u = np.ascontiguousarray(np.asarray(uv, np.float64).reshape(-1))
    cnt = QexCounts()
    args = (None, nV, nF, _ptr(V, _D), _ptr(F, _I32), N, _ptr(mt, _I32), _ptr(sg, _I32), _ptr(u, _D))
    s = L.mk_qex_extract(*args, None, 0, None, None, None, 0, None, None, 0, None, 0, None, C.byref(cnt))
    if s:
        raise MetrikoError(s, "mk_qex_extract")
    nq = cnt.n_vq + cnt.n_eq + cnt.n_fq
    uvo = np.zeros((3 * nF, 2))
    qu, qp, qs = np.zeros((nq, 2)), np.zeros((nq, 3)), np.zeros(nq, np.int32)
    pid, puv = np.zeros((cnt.n_ports, 6), np.int32), np.zeros((cnt.n_ports, 4))
    ep, fh = np.zeros((cnt.n_edges, 2), np.int32), np.zeros((cnt.n_faces, 4), np.int32)
    s = L.mk_qex_extract(*args, _ptr(uvo, _D), nq, _ptr(qu, _D), _ptr(qp, _D), _ptr(qs, _I32), cnt.n_ports,
                         _ptr(pid, _I32), _ptr(puv, _D), cnt.n_edges, _ptr(ep, _I32), cnt.n_faces, _ptr(fh, _I32),
                         C.byref(cnt))
    if s:
        raise MetrikoError(s, "mk_qex_extract")
    kind = np.repeat(np.arange(3, dtype=np.int32), [cnt.n_vq, cnt.n_eq, cnt.n_fq])
    return {"uv": uvo, "q_uv": qu, "q_pos": qp, "q_sid": qs, "q_kind": kind, "port_ids": pid, "port_uvdir": puv,
            "edges": ep, "faces": fh}


def combinatorics(V, F, compressed, N=4):
    """Host-only stages after the field solve (roots, Curl matching, singularities, seam,
    cut mesh, combed matching, IGM setup sizes); no GPU needed."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    nE = hmesh_export(V, F)["nE"]
    comp = np.ascontiguousarray(np.stack([np.real(compressed), np.imag(compressed)], 1).reshape(-1))
    sing = np.zeros(nV, np.int32)
    rawm = np.zeros(nE, np.int32)
    mat = np.zeros(nE, np.int32)
    seam = np.zeros(nE, np.uint8)
    cut = np.zeros((nF, 3), np.int32)
    sizes = np.zeros(4, np.int32)
    ext = np.zeros((nF, 12))
    s = L.mk_combinatorics(nV, nF, _ptr(V, _D), _ptr(F, _I32), N, _ptr(comp, _D), _ptr(sing, _I32),
                           _ptr(rawm, _I32), _ptr(mat, _I32), _ptr(seam, _U8), _ptr(cut, _I32), _ptr(sizes, _I32),
                           _ptr(ext, _D))
    if s:
        raise MetrikoError(s, "mk_combinatorics")
    return {"singular": sing, "raw_matching": rawm, "matching": mat, "seam": seam.astype(bool), "cut_idx": cut,
            "cutV": int(sizes[0]), "nS": int(sizes[1]), "nT": int(sizes[2]), "constraint_rows": int(sizes[3]),
            "ext": ext}


def _cplx_in(z):
    return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], 1).reshape(-1))


def curvature_dir(V, F):
    """principalCurvatureDir of the library's Hmesh (host only): complex per face."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    D = np.zeros(2 * F.shape[0])
    s = L.mk_curvature_dir(V.shape[0], F.shape[0], _ptr(V, _D), _ptr(F, _I32), _ptr(D, _D))
    if s:
        raise MetrikoError(s, "mk_curvature_dir")
    return D[0::2] + 1j * D[1::2]


def field_matching(V, F, compressed, N=4, kind="curl"):
    """computeMatching(Principal | Curl) + computeSingular on a given compressed field (host only)."""
    L = lib()
    V, F = _mesh_arrays(V, F)
    nV, nF = V.shape[0], F.shape[0]
    nE = hmesh_export(V, F)["nE"]
    comp = _cplx_in(compressed)
    sing = np.zeros(nV, np.int32)
    mat = np.zeros(nE, np.int32)
    eff = np.zeros(nE)
    s = L.mk_field_matching(nV, nF, _ptr(V, _D), _ptr(F, _I32), N, _ptr(comp, _D),
                            MATCH_CURL if kind == "curl" else MATCH_PRINCIPAL, _ptr(sing, _I32), _ptr(mat, _I32),
                            _ptr(eff, _D))
    if s:
        raise MetrikoError(s, "mk_field_matching")
    return {"singular": sing, "matching": mat, "effort": eff}


class SizingState:
    """The last LM round's system kept from a pipeline run (mk_keep_igm_state): the sized solve
    sigma -> corner uv and its adjoint (mk_igm_sized_uv / mk_igm_sized_uv_grad)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        vals = [C.c_int() for _ in range(5)]
        ctx._check(ctx.L.mk_igm_state_sizes(h, *[C.byref(v) for v in vals]), "mk_igm_state_sizes")
        self.nF, self.nfull, self.k, self.nfixed, self.has_corner_map = [v.value for v in vals]

    def uv(self, sigma):
        sig = np.ascontiguousarray(sigma, np.float64)
        assert sig.shape == (self.nF,)
        fx = np.zeros(self.nfull)
        uv = np.zeros(6 * self.nF)
        self.ctx._check(self.ctx.L.mk_igm_sized_uv(self.ctx.h, self.h, _ptr(sig, _D), _ptr(fx, _D), _ptr(uv, _D)),
                        "mk_igm_sized_uv")
        return fx, uv

    def grad(self, sigma, g_uv):
        sig = np.ascontiguousarray(sigma, np.float64)
        g = np.ascontiguousarray(g_uv, np.float64)
        assert sig.shape == (self.nF,) and g.shape == (6 * self.nF,)
        out = np.zeros(self.nF)
        self.ctx._check(self.ctx.L.mk_igm_sized_uv_grad(self.ctx.h, self.h, _ptr(sig, _D), _ptr(g, _D), _ptr(out, _D)),
                        "mk_igm_sized_uv_grad")
        return out

    def close(self):
        if self.h:
            self.ctx.L.mk_igm_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One mk_ctx (HIP stream + scratch) on one device."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        s = self.L.mk_init(C.byref(self.h), device)
        if s:
            raise MetrikoError(s, "mk_init")

    def close(self):
        if self.h:
            self.L.mk_finalize(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self):
        buf = C.create_string_buffer(1024)
        self.L.mk_last_error(self.h, buf, 1024)
        return buf.value.decode()

    def _check(self, s, what):
        if s:
            raise MetrikoError(s, what, self.last_error())

    def set_timing(self, on):
        self._check(self.L.mk_set_timing(self.h, int(on)), "mk_set_timing")

    def set_graphs(self, on):
        self._check(self.L.mk_set_graphs(self.h, int(on)), "mk_set_graphs")

    def timing_report(self, reset=True):
        """Per-kernel SpMV timings (hipEvent) since the last reset, as a list of dicts."""
        import json
        buf = C.create_string_buffer(1 << 20)
        self._check(self.L.mk_timing_report(self.h, int(reset), buf, 1 << 20), "mk_timing_report")
        return json.loads(buf.value.decode())

    # --- operators -------------------------------------------------------------------
    def op_real(self, A, o=None):
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sort_indices()
        ptr = A.indptr.astype(np.int64)
        col = A.indices.astype(np.int32)
        val = A.data.astype(np.float64)
        h = C.c_void_p()
        self._check(self.L.mk_op_create_real(self.h, A.shape[0], _ptr(ptr, _I64), _ptr(col, _I32), _ptr(val, _D),
                                             C.byref(o or opts()), C.byref(h)), "mk_op_create_real")
        return Op(self, h, A.shape[0], False)

    def op_complex(self, A, o=None):
        import scipy.sparse as sp
        A = sp.csr_matrix(A, dtype=np.complex128)
        A.sort_indices()
        ptr = A.indptr.astype(np.int64)
        col = A.indices.astype(np.int32)
        val = np.ascontiguousarray(np.stack([A.data.real, A.data.imag], 1).reshape(-1))
        h = C.c_void_p()
        self._check(self.L.mk_op_create_complex(self.h, A.shape[0], _ptr(ptr, _I64), _ptr(col, _I32), _ptr(val, _D),
                                                C.byref(o or opts()), C.byref(h)), "mk_op_create_complex")
        return Op(self, h, A.shape[0], True)

    def field_laplacian(self, face_halfs, twin, half_face, farg, N=4, o=None):
        fh = np.ascontiguousarray(face_halfs, np.int32).reshape(-1)
        tw = np.ascontiguousarray(twin, np.int32)
        hf = np.ascontiguousarray(half_face, np.int32)
        fa = np.ascontiguousarray(farg, np.float64)
        h = C.c_void_p()
        nF = fh.size // 3
        self._check(self.L.mk_field_laplacian(self.h, nF, tw.size, _ptr(fh, _I32), _ptr(tw, _I32), _ptr(hf, _I32),
                                              _ptr(fa, _D), N, C.byref(o or opts()), C.byref(h)),
                    "mk_field_laplacian")
        return Op(self, h, nF, True)

    def smallest_eig(self, Lop, mass, u0, n_iter=50, o=None):
        mass = np.ascontiguousarray(mass, np.float64)
        u = np.ascontiguousarray(np.stack([np.real(u0), np.imag(u0)], 1).reshape(-1))
        x = np.zeros_like(u)
        st = Stats()
        self._check(self.L.mk_field_smallest_eig(self.h, Lop.h, _ptr(mass, _D), _ptr(u, _D), n_iter,
                                                 C.byref(o or opts()), _ptr(x, _D), C.byref(st)),
                    "mk_field_smallest_eig")
        return x[0::2] + 1j * x[1::2], struct_dict(st)

    def field(self, V, F, N=4, kind="smoothest", u0=None, n_iter=50, o=None):
        """FaceRosyField(mesh, N, Smoothest | CurvatureAligned) (mk_field_solve): compressed field."""
        V, F = _mesh_arrays(V, F)
        nF = F.shape[0]
        u = None if u0 is None else _cplx_in(u0)
        x = np.zeros(2 * nF)
        st = Stats()
        ft = FIELD_CURVATURE_ALIGNED if kind == "curvature" else FIELD_SMOOTHEST
        self._check(self.L.mk_field_solve(self.h, V.shape[0], nF, _ptr(V, _D), _ptr(F, _I32), N, ft,
                                          None if u is None else _ptr(u, _D), n_iter, C.byref(o or opts()),
                                          _ptr(x, _D), C.byref(st)), "mk_field_solve")
        return x[0::2] + 1j * x[1::2], struct_dict(st)

    def batch(self, meshes, N=4, n_iter=50, u0s=None, o=None):
        """mk_batch_run: Smoothest field + u/v Poisson for every (V, F) in meshes in one batch.
        Returns a list of {"compressed", "u", "v"} and the batch statistics."""
        M = len(meshes)
        Vs, Fs, comps, uvs, u0a = [], [], [], [], []
        for i, (V, F) in enumerate(meshes):
            V, F = _mesh_arrays(V, F)
            Vs.append(V)
            Fs.append(F)
            comps.append(np.zeros(2 * F.shape[0]))
            uvs.append(np.zeros(2 * V.shape[0]))
            if u0s is not None:
                u0a.append(_cplx_in(u0s[i]))
        nV = (C.c_int32 * M)(*[V.shape[0] for V in Vs])
        nF = (C.c_int32 * M)(*[F.shape[0] for F in Fs])
        Vp = (_D * M)(*[_ptr(V, _D) for V in Vs])
        Fp = (_I32 * M)(*[_ptr(F, _I32) for F in Fs])
        Cp = (_D * M)(*[_ptr(x, _D) for x in comps])
        Up = (_D * M)(*[_ptr(x, _D) for x in uvs])
        U0 = (_D * M)(*[_ptr(x, _D) for x in u0a]) if u0s is not None else None
        st = BatchStats()
        self._check(self.L.mk_batch_run(self.h, M, nV, nF, Vp, Fp, N, n_iter, U0, C.byref(o or opts()), Cp, Up,
                                        C.byref(st)), "mk_batch_run")
        out = [{"compressed": x[0::2] + 1j * x[1::2], "u": y[0::2].copy(), "v": y[1::2].copy()}
               for x, y in zip(comps, uvs)]
        return out, struct_dict(st)

    def keep_igm_state(self, on=True):
        self._check(self.L.mk_keep_igm_state(self.h, int(on)), "mk_keep_igm_state")

    def take_igm_state(self):
        h = C.c_void_p()
        self._check(self.L.mk_igm_state_take(self.h, C.byref(h)), "mk_igm_state_take")
        return SizingState(self, h)

    def autotune_spmv(self, A, o=None):
        """mk_autotune_spmv on a scipy CSR matrix: the fastest layout's options and its time (us)."""
        A = A.tocsr()
        cplx = np.iscomplexobj(A.data)
        rp = np.ascontiguousarray(A.indptr, np.int64)
        ci = np.ascontiguousarray(A.indices, np.int32)
        val = _cplx_in(A.data) if cplx else np.ascontiguousarray(A.data, np.float64)
        best = SolveOpts()
        us = C.c_double()
        self._check(self.L.mk_autotune_spmv(self.h, A.shape[0], _ptr(rp, _I64), _ptr(ci, _I32), _ptr(val, _D),
                                            int(cplx), C.byref(o or opts()), C.byref(best), C.byref(us)),
                    "mk_autotune_spmv")
        return best, us.value

    def iterative_rounding(self, fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n, nF,
                           F2, seamless=True, round_seams=False, local_injectivity=False, o=None,
                           cut_idx=None, fullx2Nfunc=None, iter_cb=None):
        """iterative_rounding (iter_rounding.h:33-51) through mk_iterative_rounding: Cfull and
        G2 as scipy CSR; returns fullx, success, decisions [(index, value, margin)], stats.
        cut_idx (cut faces x 3) and fullx2Nfunc (scipy CSR) give the naive pass the
        locally_injectivity_check RosyParameterization passes as iter_cb; iter_cb(fullx) -> bool
        is a host callback in its place (mk_set_naive_iter_callback)."""
        cb_ref = None
        if iter_cb is not None:
            def _tramp(user, x, nx):
                return 1 if iter_cb(np.ctypeslib.as_array(x, shape=(nx,)).copy()) else 0
            cb_ref = NAIVE_ITER_CB(_tramp)
            self._check(self.L.mk_set_naive_iter_callback(self.h, cb_ref, None), "mk_set_naive_iter_callback")
        try:
            return self._iterative_rounding_checked(fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull,
                                                    G2, N, n, nF, F2, seamless, round_seams, local_injectivity, o,
                                                    cut_idx, fullx2Nfunc)
        finally:
            if cb_ref is not None:
                self.L.mk_set_naive_iter_callback(self.h, NAIVE_ITER_CB(), None)

    def _iterative_rounding_checked(self, fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n,
                                    nF, F2, seamless, round_seams, local_injectivity, o, cut_idx, fullx2Nfunc):
        if cut_idx is not None:
            ci = np.ascontiguousarray(cut_idx, np.int32).reshape(-1)
            f2n = fullx2Nfunc.tocsr()
            fp, fc, fv_ = (np.ascontiguousarray(f2n.indptr, np.int64), np.ascontiguousarray(f2n.indices, np.int32),
                           np.ascontiguousarray(f2n.data, np.float64))
            self._check(self.L.mk_igm_injectivity_check(self.h, N, len(ci) // 3, _ptr(ci, _I32), f2n.shape[0],
                                                         f2n.shape[1], _ptr(fp, _I64), _ptr(fc, _I32), _ptr(fv_, _D)),
                        "mk_igm_injectivity_check")
        try:
            return self._iterative_rounding(fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n,
                                            nF, F2, seamless, round_seams, local_injectivity, o)
        finally:
            if cut_idx is not None:
                self.L.mk_igm_injectivity_check(self.h, N, 0, None, 0, 0, None, None, None)

    def _iterative_rounding(self, fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n, nF,
                            F2, seamless, round_seams, local_injectivity, o):
        fi = np.ascontiguousarray(fixedIdcs, np.int32)
        fv = np.ascontiguousarray(fixedVals, np.float64)
        si = np.ascontiguousarray(singularIdcs, np.int32)
        ii = np.ascontiguousarray(integerIdcs, np.int32)
        Cf = Cfull.tocsr()
        G = G2.tocsr()
        cp, cc, cv = (np.ascontiguousarray(Cf.indptr, np.int64), np.ascontiguousarray(Cf.indices, np.int32),
                      np.ascontiguousarray(Cf.data, np.float64))
        gp, gc, gv = (np.ascontiguousarray(G.indptr, np.int64), np.ascontiguousarray(G.indices, np.int32),
                      np.ascontiguousarray(G.data, np.float64))
        f2 = np.ascontiguousarray(F2, np.float64)
        fl = IgmFlags(int(seamless), int(round_seams), int(local_injectivity), 0)
        fullx = np.zeros(G.shape[1])
        cap = len(si) + len(ii) + 1
        di = np.zeros(cap, np.int32)
        dv = np.zeros(cap, np.int32)
        dm = np.zeros(cap)
        nd = C.c_int()
        ok = C.c_int()
        st = IgmStats()
        self._check(self.L.mk_iterative_rounding(
            self.h, len(fi), _ptr(fi, _I32), _ptr(fv, _D), len(si), _ptr(si, _I32), len(ii), _ptr(ii, _I32),
            float(length), Cf.shape[0], Cf.shape[1], _ptr(cp, _I64), _ptr(cc, _I32), _ptr(cv, _D), G.shape[0],
            G.shape[1], _ptr(gp, _I64), _ptr(gc, _I32), _ptr(gv, _D), N, n, nF, C.byref(fl), _ptr(f2, _D),
            C.byref(o or opts()), _ptr(fullx, _D), C.byref(ok), cap, _ptr(di, _I32), _ptr(dv, _I32), _ptr(dm, _D),
            C.byref(nd), C.byref(st)), "mk_iterative_rounding")
        m = min(nd.value, cap)
        dec = list(zip(di[:m].tolist(), dv[:m].tolist(), dm[:m].tolist()))
        return fullx, bool(ok.value), dec, struct_dict(st)

    def pipeline(self, V, F, ratio=0.03, u0=None, n_iter=50, field_opts=None, igm_opts=None):
        """example_igm in one call (mk_pipeline_run); returns outputs and per-stage info."""
        V, F = _mesh_arrays(V, F)
        nV, nF = V.shape[0], F.shape[0]
        nE = 3 * nF  # at least the edge count (every edge has a side); the exact count comes back in info
        comp = np.zeros(2 * nF)
        sing = np.zeros(nV, np.int32)
        mat = np.zeros(nE, np.int32)
        seam = np.zeros(nE, np.uint8)
        cut = np.zeros((nF, 3), np.int32)
        uv = np.zeros((nF * 3, 2))
        cap = 2 * (nV + 2 * nE) + 64
        fullx = np.zeros(cap)
        dcap = 4 * nV + 64
        di = np.zeros(dcap, np.int32)
        dv = np.zeros(dcap, np.int32)
        u = None
        if u0 is not None:
            u = np.ascontiguousarray(np.stack([np.real(u0), np.imag(u0)], 1).reshape(-1))
        info = PipelineInfo()
        self._check(self.L.mk_pipeline_run(self.h, nV, nF, _ptr(V, _D), _ptr(F, _I32), 4, ratio, _ptr(u, _D), n_iter,
                                           C.byref(field_opts) if field_opts else None,
                                           C.byref(igm_opts) if igm_opts else None,
                                           _ptr(comp, _D), _ptr(sing, _I32), _ptr(mat, _I32), _ptr(seam, _U8),
                                           _ptr(cut, _I32), _ptr(uv, _D), cap, _ptr(fullx, _D), dcap, _ptr(di, _I32),
                                           _ptr(dv, _I32), C.byref(info)), "mk_pipeline_run")
        inf = struct_dict(info)
        ne = inf["nE"]
        return {"compressed": comp[0::2] + 1j * comp[1::2], "singular": sing, "matching": mat[:ne].copy(),
                "seam": seam[:ne].astype(bool), "cut_idx": cut, "uv": uv, "fullx": fullx[:inf["nfullx"]],
                "decisions": list(zip(di[:inf["ndec"]].tolist(), dv[:inf["ndec"]].tolist())), "info": inf}


class Op:
    def __init__(self, ctx, h, n, cplx):
        self.ctx, self.h, self.n, self.cplx = ctx, h, n, cplx

    def close(self):
        if self.h:
            self.ctx.L.mk_op_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _pack(self, x):
        if self.cplx:
            return np.ascontiguousarray(np.stack([np.real(x), np.imag(x)], 1).reshape(-1))
        return np.ascontiguousarray(x, np.float64)

    def _unpack(self, y):
        return y[0::2] + 1j * y[1::2] if self.cplx else y

    def apply(self, x):
        xp = self._pack(x)
        y = np.zeros_like(xp)
        self.ctx._check(self.ctx.L.mk_op_apply(self.ctx.h, self.h, _ptr(xp, _D), _ptr(y, _D)), "mk_op_apply")
        return self._unpack(y)

    def solve(self, b, x0=None, o=None):
        bp = self._pack(b)
        x = self._pack(x0) if x0 is not None else np.zeros_like(bp)
        x = x.copy()
        st = Stats()
        self.ctx._check(self.ctx.L.mk_spd_solve(self.ctx.h, self.h, _ptr(bp, _D), _ptr(x, _D), C.byref(o or opts()),
                                                C.byref(st)), "mk_spd_solve")
        return self._unpack(x), struct_dict(st)

    def solve_many(self, B, o=None):
        """Solves A X = B for the columns of B (n x m, real operators)."""
        B = np.asarray(B, np.float64)
        if B.ndim == 1:
            B = B[:, None]
        m = B.shape[1]
        Bc = np.ascontiguousarray(B.T)  # column-major n x m == row-major m x n
        X = np.zeros_like(Bc)
        st = Stats()
        self.ctx._check(self.ctx.L.mk_spd_solve_many(self.ctx.h, self.h, m, _ptr(Bc, _D), _ptr(X, _D),
                                                     C.byref(o or opts()), C.byref(st)), "mk_spd_solve_many")
        return X.T.copy(), struct_dict(st)

    def export_csr(self):
        import scipy.sparse as sp
        nnz = C.c_int64(0)
        self.ctx._check(self.ctx.L.mk_op_export_csr(self.ctx.h, self.h, C.byref(nnz), None, None, None), "export")
        n = nnz.value
        ptr = np.zeros(self.n + 1, np.int64)
        col = np.zeros(n, np.int32)
        vw = 2 if self.cplx else 1
        val = np.zeros(n * vw)
        self.ctx._check(self.ctx.L.mk_op_export_csr(self.ctx.h, self.h, C.byref(nnz), _ptr(ptr, _I64),
                                                    _ptr(col, _I32), _ptr(val, _D)), "export")
        v = val[0::2] + 1j * val[1::2] if self.cplx else val
        return sp.csr_matrix((v, col, ptr), shape=(self.n, self.n))
