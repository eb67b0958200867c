"""mk_spd_solve_noalloc (the allocation-free variant of mk_spd_solve, its scratch in the
library's static per-device array) against mk_spd_solve through the C ABI: the same FP64 PCG
kernels in the same order, so the solutions agree bit for bit; two calls in a row, the second
with a smaller operator (the scratch keeps the first call's contents), real and complex
operators, and the size / mode limits reported as errors."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import metriko_amd as mk

pytestmark = pytest.mark.gpu

SCRATCH_LEN = 1 << 19  # scalar unknowns the scratch holds per work vector


def lap2d(nx, shift):
    T = sp.diags([-np.ones(nx - 1), 2.0 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1])
    return (sp.kron(T, sp.identity(nx)) + sp.kron(sp.identity(nx), T) + shift * sp.identity(nx * nx)).tocsr()


def _hip():
    """The HIP runtime the library already loaded (by its soname): the caller's own device
    allocations and copies, made before and after the allocation-free call."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


H2D, D2H = 1, 2


def noalloc_solve(ctx, op, b, o):
    """b, x as device arrays owned by the caller, then the allocation-free call."""
    hip = _hip()
    bp = op._pack(b)
    nbytes = bp.size * 8
    bd, xd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(bd), nbytes) == 0 and hip.hipMalloc(C.byref(xd), nbytes) == 0
    try:
        x = np.zeros_like(bp)
        assert hip.hipMemcpy(bd, bp.ctypes.data, nbytes, H2D) == 0
        assert hip.hipMemcpy(xd, x.ctypes.data, nbytes, H2D) == 0
        st = mk.Stats()
        s = ctx.L.mk_spd_solve_noalloc(ctx.h, op.h, bd, xd, C.byref(o), C.byref(st))
        ctx._check(s, "mk_spd_solve_noalloc")
        assert hip.hipMemcpy(x.ctypes.data, xd, nbytes, D2H) == 0
        return op._unpack(x), mk.struct_dict(st)
    finally:
        hip.hipFree(bd)
        hip.hipFree(xd)


@pytest.mark.parametrize("precond", [mk.PRECOND_JACOBI, mk.PRECOND_NONE])
def test_noalloc_matches_spd_solve_twice_smaller_second(gpu_ctx, precond):
    o = mk.opts(precision=mk.PREC_FP64, precond=precond)
    rng = np.random.default_rng(precond)
    for nx in (70, 45):  # the second call on a smaller operator
        A = lap2d(nx, 1e-2)
        op = gpu_ctx.op_real(A, o)
        b = rng.standard_normal(A.shape[0])
        x_ref, st_ref = op.solve(b, o=o)
        x, st = noalloc_solve(gpu_ctx, op, b, o)
        print(f"[noalloc] real nx={nx} precond={precond}: {st['iterations']} iterations "
              f"(mk_spd_solve {st_ref['iterations']}), residual {st['rel_residual']:.2e}", flush=True)
        assert np.array_equal(x, x_ref)
        assert st["iterations"] == st_ref["iterations"]
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10
        op.close()


def test_noalloc_complex_matches_spd_solve(gpu_ctx):
    o = mk.opts(precision=mk.PREC_FP64, precond=mk.PRECOND_JACOBI)
    rng = np.random.default_rng(5)
    for nx in (40, 24):
        n = nx * nx
        S = sp.random(n, n, density=4.0 / n, random_state=nx, format="csr")
        A = (lap2d(nx, 0.5) + 0.05j * (S - S.T)).tocsr()   # Hermitian positive definite
        op = gpu_ctx.op_complex(A, o)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x_ref, _ = op.solve(b, o=o)
        x, st = noalloc_solve(gpu_ctx, op, b, o)
        print(f"[noalloc] complex nx={nx}: {st['iterations']} iterations", flush=True)
        assert np.array_equal(x, x_ref)
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10
        op.close()


def test_noalloc_rejects_beyond_limits(gpu_ctx):
    o = mk.opts(precision=mk.PREC_FP64, precond=mk.PRECOND_JACOBI)
    n = SCRATCH_LEN + 1
    A = sp.diags([-np.ones(n - 1), 2.5 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    op = gpu_ctx.op_real(A, o)
    with pytest.raises(mk.MetrikoError) as ei:
        noalloc_solve(gpu_ctx, op, np.ones(n), o)
    assert ei.value.status == 1 and "scratch limit" in str(ei.value)
    op.close()
    small = gpu_ctx.op_real(lap2d(10, 1e-2), o)
    with pytest.raises(mk.MetrikoError) as ei:   # the multigrid set-up would allocate
        noalloc_solve(gpu_ctx, small, np.ones(100), mk.opts(precision=mk.PREC_FP64, precond=mk.PRECOND_AMG))
    assert ei.value.status == 8
    with pytest.raises(mk.MetrikoError) as ei:   # the FP32 modes need FP32 work vectors
        noalloc_solve(gpu_ctx, small, np.ones(100), mk.opts(precision=mk.PREC_MIXED, precond=mk.PRECOND_JACOBI))
    assert ei.value.status == 8
    x, st = noalloc_solve(gpu_ctx, small, np.ones(100), o)   # still usable after the errors
    assert st["rel_residual"] <= 1e-10
    small.close()
