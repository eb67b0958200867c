// Sparse kernels and the PCG / refinement solvers of the metriko HIP library (gfx950).
//
// Layout: SELL-C-sigma (slice height C = 64 = one wavefront by default): entries of
// slice s are stored column-major (val[slice_ptr[s] + j*C + lane]) so every load of
// a wavefront is one contiguous, coalesced 64-lane access.  Rows may be sorted by
// length inside windows of sigma rows (perm maps a stored row to the original row).
//
// Reductions are deterministic: each block reduces with xor-butterfly shuffles
// (DPP / ds_swizzle on CDNA) and LDS, writes one partial per block, and the last
// block to arrive (atomic ticket) sums the partials in a fixed order.  Reruns are
// therefore bit-identical.
#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstring>
#include <numeric>
#include "mk_internal.h"

#include "sell.h"

namespace mk {

hipEvent_t next_event(mk_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        MK_HIP(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

void timing_flush(mk_ctx* c) {
    if (!c->time_spmv || c->ev_used == 0) { c->ev_used = 0; return; }
    MK_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0.f;
        MK_HIP(hipEventElapsedTime(&ms, c->ev_pool[i], c->ev_pool[i + 1]));
        c->spmv_ms += ms;
        const std::string& k = c->ev_key[i / 2];
        const bool cont = !k.empty() && k[0] == '+';  // second launch of the previous call
        const std::string key = cont ? k.substr(1) : k;
        auto it = std::find_if(c->kstats.begin(), c->kstats.end(), [&](const auto& p) { return p.first == key; });
        if (it != c->kstats.end()) { if (!cont) it->second.calls++; it->second.ms += ms; }
    }
    c->ev_used = 0;
    c->ev_key.clear();
}


// ---- second launch of the two-launch reductions ---------------------------------------------
// k_finish<K, F>: one block sums the K partials of the nparts blocks of the previous kernel
// (fixed order) and thread 0 applies F to the totals; F::guard skips it once the CG state is
// done (the working kernel then left no partials), skip (if set) does the same.
template <int K, typename F>
__global__ void __launch_bounds__(1024) k_finish(CgState* st, const double* __restrict__ partials, int nparts,
                                                 double* out, const int* skip) {
    if (skip && *skip) return;
    if (F::guard && st->done) return;
    double tot[K];
    partials_sum<K>(partials, nparts, tot);
    if (threadIdx.x == 0) F::apply(st, tot, out);
}
struct FinOut {  // plain total
    static constexpr bool guard = false;
    __device__ static void apply(CgState*, const double* tot, double* out) { out[0] = tot[0]; }
};
struct FinCgInit {  // rz, rr of the first residual
    static constexpr bool guard = false;
    __device__ static void apply(CgState* st, const double* tot, double*) {
        st->rz = tot[0];
        st->rr = tot[1];
        st->it = 0;
        st->status = MK_OK;
        st->done = (tot[1] <= st->tol2 * st->bb) ? 1 : 0;
        if (!(tot[0] == tot[0]) || !(tot[1] == tot[1])) { st->status = MK_ERR_NAN; st->done = 1; }
    }
};
struct FinCgInitZ {  // as FinCgInit, z from an external preconditioner (must be positive)
    static constexpr bool guard = false;
    __device__ static void apply(CgState* st, const double* tot, double* o) {
        FinCgInit::apply(st, tot, o);
        if (!st->done && !(tot[0] > 0.0)) { st->status = MK_ERR_NOT_SPD; st->done = 1; }
    }
};
struct FinCgAlpha {  // pq -> alpha
    static constexpr bool guard = true;
    __device__ static void apply(CgState* st, const double* tot, double*) {
        st->pq = tot[0];
        if (!(tot[0] > 0.0)) {
            st->status = (tot[0] == tot[0]) ? MK_ERR_NOT_SPD : MK_ERR_NAN;
            st->done = 1;
        } else {
            st->alpha = st->rz / tot[0];
        }
    }
};
struct FinCgUpdate {  // rz, rr after the x / r / z update -> beta, convergence
    static constexpr bool guard = true;
    __device__ static void apply(CgState* st, const double* tot, double*) {
        st->beta = tot[0] / st->rz;
        st->rz = tot[0];
        st->rr = tot[1];
        st->it += 1;
        if (!(tot[1] == tot[1]) || !(tot[0] == tot[0])) { st->status = MK_ERR_NAN; st->done = 1; }
        else if (tot[1] <= st->tol2 * st->bb) st->done = 1;
    }
};
struct FinCgRr {  // rr after the x / r update -> convergence
    static constexpr bool guard = true;
    __device__ static void apply(CgState* st, const double* tot, double*) {
        st->rr = tot[0];
        st->it += 1;
        if (!(tot[0] == tot[0])) { st->status = MK_ERR_NAN; st->done = 1; }
        else if (tot[0] <= st->tol2 * st->bb) st->done = 1;
    }
};
struct FinCgBeta {  // rz of the preconditioned residual -> beta
    static constexpr bool guard = true;
    __device__ static void apply(CgState* st, const double* tot, double*) {
        if (!(tot[0] > 0.0)) { st->status = (tot[0] == tot[0]) ? MK_ERR_NOT_SPD : MK_ERR_NAN; st->done = 1; }
        else { st->beta = tot[0] / st->rz; st->rz = tot[0]; }
    }
};
template <int K, typename F>
static inline void finish(mk_ctx* c, unsigned nparts, double* out = nullptr, const int* skip = nullptr) {
    hipLaunchKernelGGL((k_finish<K, F>), dim3(1), dim3(1024), 0, c->stream, c->state.p, (const double*)c->partials.p,
                       int(nparts), out, skip);
}

// One vector's element `row` (complex: two entries), loaded by spmv_row_gp's early hook.
template <typename T, bool CPLX>
struct RowElem {
    static constexpr int W = CPLX ? 2 : 1;
    const T* __restrict__ src;   // nullptr: nothing to load
    T v[W];
    __device__ __forceinline__ void operator()(int row) {
        if (!src) return;
#pragma unroll
        for (int q = 0; q < W; q++) v[q] = src[W * row + q];
    }
};

template <typename T, bool CPLX, int MODE>
__global__ void __launch_bounds__(256) sell_spmv(SellDev<T> A, const T* __restrict__ x, T* __restrict__ y,
                                               const T* __restrict__ b, double* partials, const int* skip) {
    if (skip && *skip) return;
    double contrib = 0.0;
    // reducing modes run a capped grid that strides over the virtual blocks
    for (int vb = blockIdx.x; vb < A.nvb; vb += gridDim.x) {
        int row = 0;
        T ar, ai;
        // the row's own element of b (modes 2, 3), y (4) or x (1), fetched early
        RowElem<T, CPLX> re{MODE == 2 || MODE == 3 ? b : MODE == 4 ? y : MODE == 1 ? x : nullptr};
        if (!spmv_row_gp<T, CPLX>(A, PlainGather<T, CPLX>{x}, vb, row, ar, ai, re)) continue;
        if constexpr (CPLX) {
            if constexpr (MODE == 2 || MODE == 3) { ar = re.v[0] - ar; ai = re.v[1] - ai; }
            if constexpr (MODE == 4) { ar += re.v[0]; ai += re.v[1]; }
            y[2 * row] = ar;
            y[2 * row + 1] = ai;
            if constexpr (MODE == 1) contrib += double(re.v[0]) * double(ar) + double(re.v[1]) * double(ai);
            if constexpr (MODE == 2) contrib += double(ar) * double(ar) + double(ai) * double(ai);
        } else {
            T a = ar;
            if constexpr (MODE == 2 || MODE == 3) a = re.v[0] - a;
            if constexpr (MODE == 4) a += re.v[0];
            y[row] = a;
            if constexpr (MODE == 1) contrib += double(re.v[0]) * double(a);
            if constexpr (MODE == 2) contrib += double(a) * double(a);
        }
    }
    if constexpr (MODE == 1 || MODE == 2) {
        double v[1] = {contrib};
        partials_write<1>(v, partials);
    }
}

// ---------------------------------------------------------------------------------------
// CG kernels.  Vectors are real arrays of length len = n (real) or 2n (complex);
// element e of a complex vector occupies entries 2e, 2e+1; the Jacobi scaling dinv
// is per element (real).
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) cg_init(int n, const T* __restrict__ r, const T* __restrict__ dinv,
                                             T* __restrict__ z, T* __restrict__ p, double* partials) {
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T d = dinv ? dinv[i] : T(1);
        if constexpr (CPLX) {
            const T r0 = r[2 * i], r1 = r[2 * i + 1];
            const T z0 = d * r0, z1 = d * r1;
            z[2 * i] = z0; z[2 * i + 1] = z1; p[2 * i] = z0; p[2 * i + 1] = z1;
            rz += double(r0) * z0 + double(r1) * z1;
            rr += double(r0) * r0 + double(r1) * r1;
        } else {
            const T r0 = r[i], z0 = d * r0;
            z[i] = z0; p[i] = z0;
            rz += double(r0) * z0;
            rr += double(r0) * r0;
        }
    }
    double v[2] = {rz, rr};
    partials_write<2>(v, partials);
}

// q = A p ; pq = p^H q ; alpha = rz / pq
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) cg_spmv(SellDev<T> A, const T* __restrict__ p, T* __restrict__ q,
                                             const CgState* st, double* partials) {
    if (st->done) return;
    double contrib = 0.0;
    for (int vb = blockIdx.x; vb < A.nvb; vb += gridDim.x) {
        int row = 0;
        T ar, ai;
        if (!spmv_row<T, CPLX>(A, p, vb, row, ar, ai)) continue;
        if constexpr (CPLX) {
            q[2 * row] = ar;
            q[2 * row + 1] = ai;
            contrib += double(p[2 * row]) * double(ar) + double(p[2 * row + 1]) * double(ai);
        } else {
            q[row] = ar;
            contrib += double(p[row]) * double(ar);
        }
    }
    double v[1] = {contrib};
    partials_write<1>(v, partials);
}

// x += alpha p ; r -= alpha q ; z = dinv r ; rz, rr ; beta
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) cg_update(int n, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                               const T* __restrict__ p, const T* __restrict__ q,
                                               const T* __restrict__ dinv, const CgState* st, double* partials) {
    if (st->done) return;
    const T alpha = T(st->alpha);
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T d = dinv ? dinv[i] : T(1);
        if constexpr (CPLX) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int k = 2 * i + c;
                x[k] += alpha * p[k];
                const T rk = r[k] - alpha * q[k];
                r[k] = rk;
                const T zk = d * rk;
                z[k] = zk;
                rz += double(rk) * zk;
                rr += double(rk) * rk;
            }
        } else {
            x[i] += alpha * p[i];
            const T rk = r[i] - alpha * q[i];
            r[i] = rk;
            const T zk = d * rk;
            z[i] = zk;
            rz += double(rk) * zk;
            rr += double(rk) * rk;
        }
    }
    double v[2] = {rz, rr};
    partials_write<2>(v, partials);
}

// preconditioned variant (z from an external preconditioner): init, update, beta
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) cg_init_z(int n, const T* __restrict__ r, const T* __restrict__ z,
                                               T* __restrict__ p, double* partials) {
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if constexpr (CPLX) {
            const T r0 = r[2 * i], r1 = r[2 * i + 1], z0 = z[2 * i], z1 = z[2 * i + 1];
            p[2 * i] = z0; p[2 * i + 1] = z1;
            rz += double(r0) * z0 + double(r1) * z1;
            rr += double(r0) * r0 + double(r1) * r1;
        } else {
            const T r0 = r[i], z0 = z[i];
            p[i] = z0;
            rz += double(r0) * z0;
            rr += double(r0) * r0;
        }
    }
    double v[2] = {rz, rr};
    partials_write<2>(v, partials);
}

template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) cg_update_noz(int n, T* __restrict__ x, T* __restrict__ r,
                                                   const T* __restrict__ p, const T* __restrict__ q,
                                                   const CgState* st, double* partials) {
    if (st->done) return;
    const T alpha = T(st->alpha);
    double rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int w = CPLX ? 2 : 1;
#pragma unroll
        for (int c = 0; c < w; c++) {
            const int k = w * i + c;
            x[k] += alpha * p[k];
            const T rk = r[k] - alpha * q[k];
            r[k] = rk;
            rr += double(rk) * rk;
        }
    }
    double v[1] = {rr};
    partials_write<1>(v, partials);
}

template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) cg_beta(int n, const T* __restrict__ r, const T* __restrict__ z,
                                             const CgState* st, double* partials) {
    if (st->done) return;
    double rz = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if constexpr (CPLX) rz += double(r[2 * i]) * z[2 * i] + double(r[2 * i + 1]) * z[2 * i + 1];
        else rz += double(r[i]) * z[i];
    }
    double v[1] = {rz};
    partials_write<1>(v, partials);
}

template <typename T>
__global__ void __launch_bounds__(256) cg_pupdate(size_t len, T* __restrict__ p, const T* __restrict__ z,
                                                const CgState* st) {
    if (st->done) return;
    const T beta = T(st->beta);
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) p[i] = z[i] + beta * p[i];
}

// ---------------------------------------------------------------------------------------
// FP64 PCG with an FP32 multigrid preconditioner (vectors as real arrays of len entries;
// complex interleaved gives Re(r^H z) as the plain sum).  The preconditioner runs in FP32,
// so it is not exactly symmetric: beta is the flexible (Polak-Ribiere) form
// beta = r_{k+1}.(z_{k+1} - z_k) / r_k.z_k, with z_k kept in FP64 (zold).

// x += alpha p ; r -= alpha q ; r32 = float(r) ; rr
__global__ void __launch_bounds__(256) cgm_update(size_t len, double* __restrict__ x, double* __restrict__ r,
                                                float* __restrict__ r32, const double* __restrict__ p,
                                                const double* __restrict__ q, CgState* st, double* partials) {
    if (st->done) return;
    const double alpha = st->alpha;
    double rr = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += size_t(gridDim.x) * blockDim.x) {
        x[i] += alpha * p[i];
        const double rk = r[i] - alpha * q[i];
        r[i] = rk;
        r32[i] = float(rk);
        rr += rk * rk;
    }
    double v[1] = {rr};
    partials_write<1>(v, partials);
}
__global__ void __launch_bounds__(1024) cgm_update_fin(CgState* st, const double* __restrict__ partials, int nparts) {
    if (st->done) return;
    double tot[1];
    partials_sum<1>(partials, nparts, tot);
    if (threadIdx.x == 0) {
        st->rr = tot[0];
        st->it += 1;
        if (!(tot[0] == tot[0])) { st->status = MK_ERR_NAN; st->done = 1; }
        else if (tot[0] <= st->tol2 * st->bb) st->done = 1;
    }
}

// beta of the flexible recurrence: a = r.z32, m = r.zold; beta = (a - m) / rz (0 on the first
// iteration); rz = a
__global__ void __launch_bounds__(256) cgm_beta2(size_t len, const double* __restrict__ r,
                                               const float* __restrict__ z32, const double* __restrict__ zold,
                                               CgState* st, double* partials) {
    if (st->done) return;
    double a = 0.0, m = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += size_t(gridDim.x) * blockDim.x) {
        a += r[i] * double(z32[i]);
        m += r[i] * zold[i];
    }
    double v[2] = {a, m};
    partials_write<2>(v, partials);
}
__global__ void __launch_bounds__(1024) cgm_beta2_fin(CgState* st, const double* __restrict__ partials, int nparts) {
    if (st->done) return;
    double tot[2];
    partials_sum<2>(partials, nparts, tot);
    if (threadIdx.x == 0) {
        if (!(tot[0] > 0.0)) { st->status = (tot[0] == tot[0]) ? MK_ERR_NOT_SPD : MK_ERR_NAN; st->done = 1; }
        else {
            st->beta = st->first ? 0.0 : (tot[0] - tot[1]) / st->rz;
            st->rz = tot[0];
            st->first = 0;
        }
    }
}

// p = z + beta p ; q = A z + beta q (q tracks A p without a second product) ; zold = z ;
// pq = p^H q ; alpha = rz / pq.  z is the FP32 preconditioned residual, A and p, q FP64.
template <bool CPLX>
struct Fp32Gather {
    const float* __restrict__ z;
    __device__ __forceinline__ void operator()(int c, double& xr, double& xi) const {
        if constexpr (CPLX) { xr = double(z[2 * c]); xi = double(z[2 * c + 1]); }
        else { xr = double(z[c]); xi = 0.0; }
    }
};
template <bool CPLX>
__global__ void __launch_bounds__(256) cgm_spmv(SellDev<double> A, const float* __restrict__ z32,
                                              double* __restrict__ p, double* __restrict__ q,
                                              double* __restrict__ zold, CgState* st, double* partials) {
    if (st->done) return;
    const double beta = st->beta;
    double contrib = 0.0;
    for (int vb = blockIdx.x; vb < A.nvb; vb += gridDim.x) {
        int row = 0;
        double ar, ai;
        if (!spmv_row_g<double, CPLX>(A, Fp32Gather<CPLX>{z32}, vb, row, ar, ai)) continue;
        if constexpr (CPLX) {
            const double z0 = double(z32[2 * row]), z1 = double(z32[2 * row + 1]);
            const double p0 = z0 + beta * p[2 * row], p1 = z1 + beta * p[2 * row + 1];
            const double q0 = ar + beta * q[2 * row], q1 = ai + beta * q[2 * row + 1];
            p[2 * row] = p0; p[2 * row + 1] = p1;
            q[2 * row] = q0; q[2 * row + 1] = q1;
            zold[2 * row] = z0; zold[2 * row + 1] = z1;
            contrib += p0 * q0 + p1 * q1;
        } else {
            const double z0 = double(z32[row]);
            const double p0 = z0 + beta * p[row], q0 = ar + beta * q[row];
            p[row] = p0;
            q[row] = q0;
            zold[row] = z0;
            contrib += p0 * q0;
        }
    }
    double v[1] = {contrib};
    partials_write<1>(v, partials);
}
__global__ void __launch_bounds__(1024) cgm_spmv_fin(CgState* st, const double* __restrict__ partials, int nparts) {
    if (st->done) return;
    double tot[1];
    partials_sum<1>(partials, nparts, tot);
    if (threadIdx.x == 0) {
        st->pq = tot[0];
        if (!(tot[0] > 0.0)) {
            st->status = (tot[0] == tot[0]) ? MK_ERR_NOT_SPD : MK_ERR_NAN;
            st->done = 1;
        } else {
            st->alpha = st->rz / tot[0];
        }
    }
}

// p = q = zold = 0 ; rr = r.r ; state for the first iteration
__global__ void __launch_bounds__(256) cgm_init2(size_t len, const double* __restrict__ r, double* __restrict__ p,
                                               double* __restrict__ q, double* __restrict__ zold, CgState* st,
                                               double* partials) {
    double rr = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += size_t(gridDim.x) * blockDim.x) {
        p[i] = 0.0;
        q[i] = 0.0;
        zold[i] = 0.0;
        rr += r[i] * r[i];
    }
    double v[1] = {rr};
    partials_write<1>(v, partials);
}
__global__ void __launch_bounds__(1024) cgm_init2_fin(CgState* st, const double* __restrict__ partials, int nparts) {
    double tot[1];
    partials_sum<1>(partials, nparts, tot);
    if (threadIdx.x == 0) {
        st->rr = tot[0];
        st->rz = 1.0;
        st->it = 0;
        st->first = 1;
        st->status = MK_OK;
        st->done = (tot[0] <= st->tol2 * st->bb) ? 1 : 0;
        if (!(tot[0] == tot[0])) { st->status = MK_ERR_NAN; st->done = 1; }
    }
}

__global__ void k_to32(size_t len, const double* __restrict__ x, float* __restrict__ y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) y[i] = float(x[i]);
}

// ---------------------------------------------------------------------------------------
// misc vector kernels
__global__ void k_axpby(size_t len, double a, const double* x, double b, double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) y[i] = a * x[i] + b * y[i];
}
__global__ void k_scale_dev(size_t len, const double* s, int invert_sqrt, double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const double f = invert_sqrt ? 1.0 / sqrt(fabs(s[0])) : s[0];
    if (i < len) y[i] *= f;
}
__global__ void k_mass_mul(int n, const double* mass, const double* u, double* b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { b[2 * i] = mass[i] * u[2 * i]; b[2 * i + 1] = mass[i] * u[2 * i + 1]; }
}
__global__ void k_mass_dot(int n, const double* mass, const double* x, double* partials) {
    double c = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += mass[i] * (x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1]);
    double v[1] = {c};
    partials_write<1>(v, partials);
}
__global__ void k_dot(size_t len, const double* x, const double* y, double* partials) {
    double c = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += size_t(gridDim.x) * blockDim.x)
        c += x[i] * y[i];
    double v[1] = {c};
    partials_write<1>(v, partials);
}
__global__ void k_cvt_scaled(size_t len, const double* x, const double* s, float* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const double f = 1.0 / sqrt(s[0]);
    if (i < len) y[i] = float(x[i] * f);
}
__global__ void k_add_scaled(size_t len, const float* d, const double* s, double* x) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const double f = sqrt(s[0]);
    if (i < len) x[i] += f * double(d[i]);
}
__global__ void k_set_state(CgState* st, double bb, double tol2) {
    st->bb = bb;
    st->tol2 = tol2;
}
__global__ void k_set_state_dev(CgState* st, const double* bb, double tol2) {
    st->bb = bb[0];
    st->tol2 = tol2;
}

// ---------------------------------------------------------------------------------------
// host side


// SELL fill: thread (slice s, lane l) writes row order[s C + l] (or s C + l) as column l of
// its slice, entry j at slice_ptr[s] + j C + l; past the row's end a padding entry (its own
// column when that is a valid column, value 0).  VW values per entry (2: complex).  With a
// renumbering the stored row `row` is CSR row rsrc[row] and a CSR column c is stored as
// cmap[c]; a row keeps its entries in CSR order.
template <int VW>
__global__ void __launch_bounds__(256) k_sell_fill(int ns, int C, int n_rows, int m, const int64_t* __restrict__ sptr,
                                                   const int* __restrict__ sw, const int* __restrict__ order,
                                                   const int64_t* __restrict__ rowptr, const int* __restrict__ col,
                                                   const double* __restrict__ val, double scale, int* __restrict__ ocol,
                                                   double* __restrict__ ov64, float* __restrict__ ov32,
                                                   const int* __restrict__ rsrc, const int* __restrict__ cmap) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(ns) * C) return;
    const int s = int(t / C), l = int(t - int64_t(s) * C);
    const int r = s * C + l;
    const int row = r < n_rows ? (order ? order[r] : r) : -1;
    const int srow = row >= 0 && rsrc ? rsrc[row] : row;
    const int64_t b = srow >= 0 ? rowptr[srow] : 0;
    const int64_t len = srow >= 0 ? rowptr[srow + 1] - b : 0;
    const int w = sw[s];
    for (int j = 0; j < w; j++) {
        const int64_t e = sptr[s] + int64_t(j) * C + l;
        if (j < len) {
            ocol[e] = cmap ? cmap[col[b + j]] : col[b + j];
#pragma unroll
            for (int q = 0; q < VW; q++) {
                const double v = scale * val[(b + j) * VW + q];
                ov64[e * VW + q] = v;
                ov32[e * VW + q] = float(v);
            }
        } else {
            ocol[e] = (row >= 0 && row < m) ? row : 0;
#pragma unroll
            for (int q = 0; q < VW; q++) { ov64[e * VW + q] = 0.0; ov32[e * VW + q] = 0.0f; }
        }
    }
}

void build_sell(mk_op* op, const int64_t* rowptr, const int32_t* col, const double* val, int C, int sigma, double scale,
                const int* rsrc, const int* cmap) {
    const int n = op->n;
    // stored row i is CSR row src(i); a CSR column c is stored as cm(c)
    auto src = [rsrc](int i) { return rsrc ? rsrc[i] : i; };
    auto cm = [cmap](int c) { return cmap ? cmap[c] : c; };
    const int vw = op->cplx ? 2 : 1;
    MK_CHECK(C > 0 && C % 64 == 0, MK_ERR_ARG, "slice height must be a positive multiple of 64");
    // every SpMV kernel is compiled with __launch_bounds__(256)
    MK_CHECK(op->block % 64 == 0 && op->block >= 64 && op->block <= 256, MK_ERR_ARG,
             "block_size must be 64, 128 or 256");
    MK_CHECK(sigma >= 1, MK_ERR_ARG, "sigma must be >= 1");
    const auto t_start = std::chrono::steady_clock::now();
    auto lap_ms = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    };
    // long rows (> max(64, 4 x median length)) go to the wavefront-per-row part
    metriko::HVec<int> lens(static_cast<size_t>(n));
#pragma omp parallel for schedule(static, 65536) if (n > 100000)
    for (int i = 0; i < n; i++) lens[i] = int(rowptr[src(i) + 1] - rowptr[src(i)]);
    // median row length (the element at sorted position n/2) from a length histogram
    int med = 0;
    if (n > 0) {
        constexpr int kHist = 4096;
        std::vector<int64_t> hist(kHist + 1, 0);
#pragma omp parallel if (n > 100000)
        {
            std::vector<int64_t> h(kHist + 1, 0);
#pragma omp for schedule(static, 65536) nowait
            for (int i = 0; i < n; i++) h[size_t(std::min(lens[i], kHist))]++;
#pragma omp critical
            for (int b = 0; b <= kHist; b++) hist[size_t(b)] += h[size_t(b)];
        }
        int64_t cum = 0;
        for (med = 0; med < kHist; med++) {
            cum += hist[size_t(med)];
            if (cum > n / 2) break;
        }
    }
    const int thr = std::max(64, 4 * med);
    // medium rows (more than kMidLen entries) get a group of lanes each instead of one lane;
    // the three classes keep ascending row order (chunked counts, then a parallel fill)
    constexpr int kMidLen = kSellMaxLen;
    auto cls = [&](int i) { return lens[i] > thr ? 2 : lens[i] > kMidLen ? 1 : 0; };
    constexpr int kChunk = 65536;
    const int nch = (n + kChunk - 1) / kChunk;
    std::vector<int> ccount(size_t(nch) * 3, 0);
#pragma omp parallel for schedule(static) if (n > 100000)
    for (int ch = 0; ch < nch; ch++)
        for (int i = ch * kChunk; i < std::min(n, (ch + 1) * kChunk); i++) ccount[size_t(ch) * 3 + cls(i)]++;
    std::vector<int> coff(size_t(nch) * 3, 0);
    int tot3[3] = {0, 0, 0};
    for (int k3 = 0; k3 < 3; k3++)
        for (int ch = 0; ch < nch; ch++) { coff[size_t(ch) * 3 + k3] = tot3[k3]; tot3[k3] += ccount[size_t(ch) * 3 + k3]; }
    metriko::HVec<int> order(static_cast<size_t>(tot3[0]));
    std::vector<int> mrows(static_cast<size_t>(tot3[1])), lrows(static_cast<size_t>(tot3[2]));
#pragma omp parallel for schedule(static) if (n > 100000)
    for (int ch = 0; ch < nch; ch++) {
        int w[3] = {coff[size_t(ch) * 3], coff[size_t(ch) * 3 + 1], coff[size_t(ch) * 3 + 2]};
        for (int i = ch * kChunk; i < std::min(n, (ch + 1) * kChunk); i++) {
            const int k3 = cls(i);
            int* dst = k3 == 0 ? order.data() : k3 == 1 ? mrows.data() : lrows.data();
            dst[w[k3]++] = i;
        }
    }
    const int ns_rows = int(order.size());
    // remaining rows of uneven length: sort inside windows of 1024 rows
    if (sigma <= 1 && ns_rows > 0) {
        int mx = 0;
        int64_t tot = 0;
        for (int r : order) { mx = std::max(mx, lens[r]); tot += lens[r]; }
        if (double(mx) > 2.0 * double(tot) / ns_rows + 8.0) sigma = 1024;
    }
    op->C = C;
    op->sigma = sigma;
    if (sigma > 1) {
#pragma omp parallel for schedule(dynamic, 16) if (ns_rows > 100000)
        for (int w0 = 0; w0 < ns_rows; w0 += sigma) {
            const int w1 = std::min(ns_rows, w0 + sigma);
            std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int a, int b) { return lens[a] > lens[b]; });
        }
    }
    bool permuted = !lrows.empty() || !mrows.empty();
    for (int i = 0; i < ns_rows && !permuted; i++) if (order[i] != i) permuted = true;
    const int n_rows = ns_rows;
    const int ns = (n_rows + C - 1) / C;
    std::vector<int64_t> sptr(size_t(ns) + 1, 0);
    std::vector<int> sw(size_t(ns), 0);
#pragma omp parallel for schedule(static, 1024) if (ns > 10000)
    for (int s = 0; s < ns; s++) {
        int w = 0;
        for (int l = 0; l < C && s * C + l < n_rows; l++) w = std::max<int>(w, lens[order[s * C + l]]);
        sw[s] = w;
    }
    for (int s = 0; s < ns; s++) sptr[s + 1] = sptr[s] + int64_t(sw[s]) * C;
    const int64_t stored = sptr[ns];
    const double t_layout = lap_ms();
    // inverse diagonal (real part) of a square operator; a rectangular one has none
    const bool square = op->m == n;
    metriko::HVec<double> dinv(square ? size_t(n) : 0);
    metriko::HVec<float> dinv32(square ? size_t(n) : 0);
#pragma omp parallel for schedule(static, 4096) if (n > 100000 && square)
    for (int i = 0; i < (square ? n : 0); i++) {
        double d = 0.0;
        const int si = src(i);
        for (int64_t k = rowptr[si]; k < rowptr[si + 1]; k++) if (cm(col[k]) == i) d += scale * val[k * vw];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
        dinv32[i] = float(dinv[i]);
    }
    // medium-row part (CSR, mid_lanes lanes per row: about 4 entries per lane)
    const int nmid = int(mrows.size());
    int mlanes = 4;
    if (nmid) {
        std::vector<int> ml;
        for (int r : mrows) ml.push_back(lens[r]);
        std::nth_element(ml.begin(), ml.begin() + nmid / 2, ml.end());
        const int mm = ml[size_t(nmid / 2)];
        while (mlanes < 16 && mlanes * 4 < mm) mlanes *= 2;
    }
    // medium and long rows as CSR, in their class order (offsets first, then a parallel fill)
    auto fill_csr = [&](const std::vector<int>& rows, std::vector<int64_t>& rp, metriko::HVec<int>& cc,
                        metriko::HVec<double>& cv, metriko::HVec<float>& cv32) {
        const int nr = int(rows.size());
        rp.assign(size_t(nr) + 1, 0);
        for (int w = 0; w < nr; w++) rp[size_t(w) + 1] = rp[size_t(w)] + lens[rows[size_t(w)]];
        const int64_t nz = rp[size_t(nr)];
        cc.resize(size_t(nz));
        cv.resize(size_t(nz) * vw);
        cv32.resize(size_t(nz) * vw);
#pragma omp parallel for schedule(static, 256) if (nz > 200000)
        for (int w = 0; w < nr; w++) {
            const int row = src(rows[size_t(w)]);
            int64_t o = rp[size_t(w)];
            for (int64_t k = rowptr[row]; k < rowptr[row + 1]; k++, o++) {
                cc[size_t(o)] = cm(col[k]);
                for (int c = 0; c < vw; c++) {
                    const double v = scale * val[k * vw + c];
                    cv[size_t(o) * vw + c] = v;
                    cv32[size_t(o) * vw + c] = float(v);
                }
            }
        }
    };
    std::vector<int64_t> mptr;
    metriko::HVec<int> mcol;
    metriko::HVec<double> mval;
    metriko::HVec<float> mval32;
    fill_csr(mrows, mptr, mcol, mval, mval32);
    // long-row part (CSR)
    const int nlong = int(lrows.size());
    std::vector<int64_t> lptr;
    metriko::HVec<int> lcol;
    metriko::HVec<double> lval;
    metriko::HVec<float> lval32;
    fill_csr(lrows, lptr, lcol, lval, lval32);
    // keep every uploaded array non-empty (an operator may consist of long rows only)
    if (sw.empty()) sw.push_back(0);
    if (order.empty()) order.push_back(0);
    const double t_fill = lap_ms();
    hipStream_t st = op->ctx->stream;
    op->nslices = ns;
    op->stored = stored;  // SELL entries only (the long part is long_ptr[nlong] entries)
    op->n_short = n_rows;
    op->nlong = nlong;
    op->slice_ptr.upload(sptr.data(), sptr.size(), st);
    op->slice_w.upload(sw.data(), std::max<size_t>(sw.size(), 1), st);
    const size_t nst = size_t(std::max<int64_t>(stored, 1));
    if (square) {
        op->dinv64.upload(dinv.data(), dinv.size(), st);
        op->dinv32.upload(dinv32.data(), dinv32.size(), st);
    } else {
        op->dinv64.release();
        op->dinv32.release();
    }
    if (permuted) op->perm.upload(order.data(), std::max<size_t>(order.size(), 1), st);
    else op->perm.release();
    // the SELL arrays are filled on the device from the CSR arrays (one upload of the CSR
    // instead of a host fill and an upload of three padded arrays)
    op->col.alloc(nst);
    op->val64.alloc(nst * vw);
    op->val32.alloc(nst * vw);
    if (stored > 0) {
        const int64_t nnz_all = rowptr[n];
        DevBuf<int64_t> drp;
        DevBuf<int> dcol, drs, dcm;
        DevBuf<double> dval;
        if (rsrc) drs.upload(rsrc, size_t(n), st);
        if (cmap) dcm.upload(cmap, size_t(op->m), st);
        drp.upload(rowptr, size_t(n) + 1, st);
        dcol.upload(col, size_t(std::max<int64_t>(nnz_all, 1)), st);
        dval.upload(val, size_t(std::max<int64_t>(nnz_all, 1)) * vw, st);
        const size_t nthr = size_t(ns) * size_t(C);
        const unsigned g = unsigned((nthr + 255) / 256);
        if (vw == 2)
            hipLaunchKernelGGL((k_sell_fill<2>), dim3(g), dim3(256), 0, st, ns, C, n_rows, op->m, op->slice_ptr.p,
                               op->slice_w.p, (const int*)(permuted ? op->perm.p : nullptr), (const int64_t*)drp.p,
                               (const int*)dcol.p, (const double*)dval.p, scale, op->col.p, op->val64.p, op->val32.p,
                               (const int*)drs.p, (const int*)dcm.p);
        else
            hipLaunchKernelGGL((k_sell_fill<1>), dim3(g), dim3(256), 0, st, ns, C, n_rows, op->m, op->slice_ptr.p,
                               op->slice_w.p, (const int*)(permuted ? op->perm.p : nullptr), (const int64_t*)drp.p,
                               (const int*)dcol.p, (const double*)dval.p, scale, op->col.p, op->val64.p, op->val32.p,
                               (const int*)drs.p, (const int*)dcm.p);
        MK_HIP(hipGetLastError());
        MK_HIP(hipStreamSynchronize(st));  // the temporaries are freed on return
    } else {
        MK_HIP(hipMemsetAsync(op->col.p, 0, nst * sizeof(int), st));
        MK_HIP(hipMemsetAsync(op->val64.p, 0, nst * vw * sizeof(double), st));
        MK_HIP(hipMemsetAsync(op->val32.p, 0, nst * vw * sizeof(float), st));
    }
    op->nmid = nmid;
    op->mid_lanes = mlanes;
    if (nmid) {
        op->mid_rows.upload(mrows.data(), mrows.size(), st);
        op->mid_ptr.upload(mptr.data(), mptr.size(), st);
        op->mid_col.upload(mcol.data(), mcol.size(), st);
        op->mid_val64.upload(mval.data(), mval.size(), st);
        op->mid_val32.upload(mval32.data(), mval32.size(), st);
    }
    if (nlong) {
        op->long_rows.upload(lrows.data(), lrows.size(), st);
        op->long_ptr.upload(lptr.data(), lptr.size(), st);
        op->long_col.upload(lcol.data(), lcol.size(), st);
        op->long_val64.upload(lval.data(), lval.size(), st);
        op->long_val32.upload(lval32.data(), lval32.size(), st);
        std::vector<int64_t> sb, se;
        std::vector<int> lsp(size_t(nlong) + 1, 0);
        for (int w = 0; w < nlong; w++) {
            for (int64_t k = lptr[w]; k < lptr[w + 1]; k += kLongSeg) {
                sb.push_back(k);
                se.push_back(std::min<int64_t>(k + kLongSeg, lptr[w + 1]));
            }
            lsp[size_t(w) + 1] = int(sb.size());
        }
        op->nseg = int(sb.size());
        op->seg_beg.upload(sb.data(), std::max<size_t>(sb.size(), 1), st);
        op->seg_end.upload(se.data(), std::max<size_t>(se.size(), 1), st);
        op->lseg_ptr.upload(lsp.data(), lsp.size(), st);
        op->seg_part.alloc(std::max<size_t>(sb.size(), 1) * 64);
    } else {
        op->nseg = 0;
    }
    MK_HIP(hipStreamSynchronize(st));
    if (n > 1000000)
        MK_LOG("[sell] %s %d x %d, %lld stored: layout %.0f ms, fill %.0f ms, upload %.0f ms\n", op->label.c_str(), n,
               op->m, (long long)stored, t_layout, t_fill - t_layout, lap_ms() - t_fill);
    const double sv64 = 8.0 * vw, sv32 = 4.0 * vw;
    const double meta = double(ns) * 12.0 + (permuted ? 4.0 * n : 0.0) + double(nlong + nmid) * 12.0;
    op->bytes64 = double(op->nnz) * (sv64 + 4.0) + meta + double(n) * sv64 * 2.0;
    op->bytes32 = double(op->nnz) * (sv32 + 4.0) + meta + double(n) * sv32 * 2.0;
}


template <typename T>
void spmv_launch(mk_ctx* c, const mk_op* A, const T* val, const T* x, T* y, const T* b, SpmvMode mode, double* out,
                 double bytes, const int* skip) {
    unsigned g = 0;
    const SellDev<T> D = sell_dev<T>(A, val, A->block, &g);
    MK_CHECK(g <= kMaxParts, MK_ERR_ARG, "operator too large for the reduction buffer");
    static const char* kname[] = {"spmv", "spmv_dot", "spmv_resid", "spmv_resid_nored", "spmv_accum"};
    SpmvTimer tm(c, bytes, A, kname[int(mode)], sizeof(T) == 4);
    const unsigned gr = std::max(1u, std::min(g, fin_grid()));  // reducing modes: capped grid
#define MK_SPMV(CPLX, MODE)                                                                                \
    MK_TLAUNCH(tm, 0, (sell_spmv<T, CPLX, MODE>), dim3((MODE == 1 || MODE == 2) ? gr : g), dim3(A->block), 0, \
               c->stream, D, x, y, b, c->partials.p, skip)
#define MK_SPMV_ALL(CPLX)                            \
    switch (mode) {                                  \
        case SPMV_PLAIN: MK_SPMV(CPLX, 0); break;    \
        case SPMV_DOT: MK_SPMV(CPLX, 1); break;      \
        case SPMV_RESID: MK_SPMV(CPLX, 2); break;    \
        case SPMV_RESID_NORED: MK_SPMV(CPLX, 3); break; \
        case SPMV_ACCUM: MK_SPMV(CPLX, 4); break;    \
    }
    if (A->cplx) { MK_SPMV_ALL(true) } else { MK_SPMV_ALL(false) }
#undef MK_SPMV_ALL
#undef MK_SPMV
    if (mode == SPMV_DOT || mode == SPMV_RESID) finish<1, FinOut>(c, gr, out, skip);
    MK_HIP(hipGetLastError());
}
// One Chebyshev step fused with the residual it needs, row-local:
//   r = b - A x_in ; d = a d + s D^-1 r (first step: d = s D^-1 r) ; x_out = x_in + d.
// The row-local operands of a Chebyshev step (element row of b, x_in, d and the inverse
// diagonal), loaded while the row's gathers are in flight.
template <typename T, bool CPLX>
struct RowOperands {
    static constexpr int W = CPLX ? 2 : 1;
    const T* __restrict__ b;
    const T* __restrict__ x;
    const T* __restrict__ d;   // nullptr: not needed
    const T* __restrict__ dinv;
    T bv[W], xv[W], dv[W], di;
    __device__ __forceinline__ void operator()(int row) {
        di = dinv[row];
#pragma unroll
        for (int q = 0; q < W; q++) {
            bv[q] = b[W * row + q];
            if (x) xv[q] = x[W * row + q];
            dv[q] = d ? d[W * row + q] : T(0);
        }
    }
};

template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) sell_cheb(SellDev<T> A, const T* __restrict__ x_in, T* __restrict__ x_out,
                                               const T* __restrict__ b, T* __restrict__ d,
                                               const T* __restrict__ dinv, T a, T s, int first, const int* skip) {
    if (skip && *skip) return;
    int row = 0;
    T ar, ai;
    RowOperands<T, CPLX> ro{b, x_in, first ? nullptr : d, dinv};
    if (!spmv_row_gp<T, CPLX>(A, PlainGather<T, CPLX>{x_in}, blockIdx.x, row, ar, ai, ro)) return;
    const T sd = s * ro.di;
    if constexpr (CPLX) {
        T d0 = sd * (ro.bv[0] - ar), d1 = sd * (ro.bv[1] - ai);
        if (!first) { d0 += a * ro.dv[0]; d1 += a * ro.dv[1]; }
        d[2 * row] = d0;
        d[2 * row + 1] = d1;
        x_out[2 * row] = ro.xv[0] + d0;
        x_out[2 * row + 1] = ro.xv[1] + d1;
    } else {
        T d0 = sd * (ro.bv[0] - ar);
        if (!first) d0 += a * ro.dv[0];
        d[row] = d0;
        x_out[row] = ro.xv[0] + d0;
    }
}

template <typename T>
void cheb_launch(mk_ctx* c, const mk_op* A, const T* x_in, T* x_out, const T* b, T* d, T a, T s, int first,
                 const int* skip) {
    unsigned g = 0;
    const T* val;
    const T* dinv;
    if constexpr (std::is_same<T, double>::value) { val = A->val64.p; dinv = A->dinv64.p; }
    else { val = A->val32.p; dinv = A->dinv32.p; }
    const SellDev<T> D = sell_dev<T>(A, val, A->block, &g);
    const double vb = double(sizeof(T)) * (A->cplx ? 2 : 1);
    const double bytes = (sizeof(T) == 8 ? A->bytes64 : A->bytes32) + double(A->n) * (vb * (first ? 2.0 : 3.0) +
                                                                                      double(sizeof(T)));
    SpmvTimer tm(c, bytes, A, "cheb", sizeof(T) == 4);
    if (A->cplx)
        MK_TLAUNCH(tm, 0, (sell_cheb<T, true>), dim3(g), dim3(A->block), 0, c->stream, D, x_in, x_out, b, d, dinv,
                   a, s, first, skip);
    else
        MK_TLAUNCH(tm, 0, (sell_cheb<T, false>), dim3(g), dim3(A->block), 0, c->stream, D, x_in, x_out, b, d, dinv,
                   a, s, first, skip);
    MK_HIP(hipGetLastError());
}
template void cheb_launch<double>(mk_ctx*, const mk_op*, const double*, double*, const double*, double*, double,
                                  double, int, const int*);
template void cheb_launch<float>(mk_ctx*, const mk_op*, const float*, float*, const float*, float*, float, float,
                                 int, const int*);

// First two Chebyshev steps from a zero start in one launch: x1 = c0 D^-1 b is formed on the
// fly inside the product (x1 never stored), then r = b - A x1, d = a x1 + s D^-1 r,
// x_out = x1 + d.
template <typename T, bool CPLX>
struct ScaledGather {
    const T* __restrict__ b;
    const T* __restrict__ dinv;
    T c0;
    __device__ __forceinline__ void operator()(int c, T& xr, T& xi) const {
        const T f = c0 * dinv[c];
        if constexpr (CPLX) { xr = f * b[2 * c]; xi = f * b[2 * c + 1]; }
        else { xr = f * b[c]; xi = T(0); }
    }
};
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) sell_cheb0(SellDev<T> A, const T* __restrict__ b, const T* __restrict__ dinv,
                                                T* __restrict__ x_out, T* __restrict__ d, T c0, T a, T s,
                                                const int* skip) {
    if (skip && *skip) return;
    int row = 0;
    T ar, ai;
    RowOperands<T, CPLX> ro{b, nullptr, nullptr, dinv};
    if (!spmv_row_gp<T, CPLX>(A, ScaledGather<T, CPLX>{b, dinv, c0}, blockIdx.x, row, ar, ai, ro)) return;
    const T di = ro.di, f = c0 * di, sd = s * di;
    if constexpr (CPLX) {
        const T x0 = f * ro.bv[0], x1 = f * ro.bv[1];
        const T d0 = a * x0 + sd * (ro.bv[0] - ar), d1 = a * x1 + sd * (ro.bv[1] - ai);
        d[2 * row] = d0;
        d[2 * row + 1] = d1;
        x_out[2 * row] = x0 + d0;
        x_out[2 * row + 1] = x1 + d1;
    } else {
        const T x0 = f * ro.bv[0];
        const T d0 = a * x0 + sd * (ro.bv[0] - ar);
        d[row] = d0;
        x_out[row] = x0 + d0;
    }
}

template <typename T>
void cheb0_launch(mk_ctx* c, const mk_op* A, const T* b, T* x_out, T* d, T c0, T a, T s, const int* skip) {
    unsigned g = 0;
    const T* val;
    const T* dinv;
    if constexpr (std::is_same<T, double>::value) { val = A->val64.p; dinv = A->dinv64.p; }
    else { val = A->val32.p; dinv = A->dinv32.p; }
    const SellDev<T> D = sell_dev<T>(A, val, A->block, &g);
    const double vb = double(sizeof(T)) * (A->cplx ? 2 : 1);
    const double bytes = (sizeof(T) == 8 ? A->bytes64 : A->bytes32) + double(A->n) * (vb + double(sizeof(T)));
    SpmvTimer tm(c, bytes, A, "cheb0", sizeof(T) == 4);
    if (A->cplx)
        MK_TLAUNCH(tm, 0, (sell_cheb0<T, true>), dim3(g), dim3(A->block), 0, c->stream, D, b, dinv, x_out, d, c0,
                   a, s, skip);
    else
        MK_TLAUNCH(tm, 0, (sell_cheb0<T, false>), dim3(g), dim3(A->block), 0, c->stream, D, b, dinv, x_out, d, c0,
                   a, s, skip);
    MK_HIP(hipGetLastError());
}
template void cheb0_launch<double>(mk_ctx*, const mk_op*, const double*, double*, double*, double, double, double,
                                  const int*);
template void cheb0_launch<float>(mk_ctx*, const mk_op*, const float*, float*, float*, float, float, float,
                                 const int*);

// y = M [u; v] for an operator whose columns [0, m1) act on u and [m1, m) on v (the
// multigrid restriction of a residual: M = [R | -R A], y = R (b - A x) in one launch)
template <typename T, bool CPLX>
struct TwoSourceGather {
    const T* __restrict__ u;
    const T* __restrict__ v;
    int m1;
    __device__ __forceinline__ void operator()(int c, T& xr, T& xi) const {
        const T* s = c < m1 ? u : v;
        const int k = c < m1 ? c : c - m1;
        if constexpr (CPLX) { xr = s[2 * k]; xi = s[2 * k + 1]; }
        else { xr = s[k]; xi = T(0); }
    }
};
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) sell_spmv2(SellDev<T> A, int m1, const T* __restrict__ u, const T* __restrict__ v,
                                                T* __restrict__ y, const int* skip) {
    if (skip && *skip) return;
    int row = 0;
    T ar, ai;
    if (!spmv_row_g<T, CPLX>(A, TwoSourceGather<T, CPLX>{u, v, m1}, blockIdx.x, row, ar, ai)) return;
    if constexpr (CPLX) { y[2 * row] = ar; y[2 * row + 1] = ai; }
    else y[row] = ar;
}

template <typename T>
void spmv2_launch(mk_ctx* c, const mk_op* A, int m1, const T* u, const T* v, T* y, const int* skip) {
    unsigned g = 0;
    const T* val;
    if constexpr (std::is_same<T, double>::value) val = A->val64.p;
    else val = A->val32.p;
    const SellDev<T> D = sell_dev<T>(A, val, A->block, &g);
    SpmvTimer tm(c, sizeof(T) == 8 ? A->bytes64 : A->bytes32, A, "restrict_resid", sizeof(T) == 4);
    if (A->cplx)
        MK_TLAUNCH(tm, 0, (sell_spmv2<T, true>), dim3(g), dim3(A->block), 0, c->stream, D, m1, u, v, y, skip);
    else
        MK_TLAUNCH(tm, 0, (sell_spmv2<T, false>), dim3(g), dim3(A->block), 0, c->stream, D, m1, u, v, y, skip);
    MK_HIP(hipGetLastError());
}
template void spmv2_launch<double>(mk_ctx*, const mk_op*, int, const double*, const double*, double*, const int*);
template void spmv2_launch<float>(mk_ctx*, const mk_op*, int, const float*, const float*, float*, const int*);

template <typename T>
__global__ void k_copy(size_t len, const T* __restrict__ x, T* __restrict__ y, const int* skip) {
    if (skip && *skip) return;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) y[i] = x[i];
}
template <typename T>
void copy_launch(mk_ctx* c, size_t len, const T* x, T* y, const int* skip) {
    hipLaunchKernelGGL(k_copy<T>, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, x, y, skip);
}
template void copy_launch<double>(mk_ctx*, size_t, const double*, double*, const int*);
template void copy_launch<float>(mk_ctx*, size_t, const float*, float*, const int*);

template void spmv_launch<double>(mk_ctx*, const mk_op*, const double*, const double*, double*, const double*,
                                  SpmvMode, double*, double, const int*);
template void spmv_launch<float>(mk_ctx*, const mk_op*, const float*, const float*, float*, const float*, SpmvMode,
                                 double*, double, const int*);

void spmv64(mk_ctx* c, const mk_op* A, const double* x, double* y, const double* b, SpmvMode mode, double* out,
            int*) {
    const double extra = (mode == SPMV_RESID || mode == SPMV_RESID_NORED || mode == SPMV_ACCUM)
                             ? A->n * 8.0 * (A->cplx ? 2 : 1) : 0.0;
    spmv_launch<double>(c, A, A->val64.p, x, y, b, mode, out, A->bytes64 + extra, nullptr);
}
void spmv32(mk_ctx* c, const mk_op* A, const float* x, float* y, SpmvMode mode, double* out, int*) {
    spmv_launch<float>(c, A, A->val32.p, x, y, nullptr, mode, out, A->bytes32, nullptr);
}

void axpby64(mk_ctx* c, size_t len, double a, const double* x, double b, double* y) {
    hipLaunchKernelGGL(k_axpby, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, a, x, b, y);
    MK_HIP(hipGetLastError());
}
void scale_dev64(mk_ctx* c, size_t len, const double* s_dev, int invert_sqrt, double* y) {
    hipLaunchKernelGGL(k_scale_dev, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, s_dev, invert_sqrt, y);
    MK_HIP(hipGetLastError());
}
void dot64(mk_ctx* c, size_t len, const double* x, const double* y, double* out_dev) {
    const unsigned g = fgrid(len);
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(256), 0, c->stream, len, x, y, c->partials.p);
    finish<1, FinOut>(c, g, out_dev);
    MK_HIP(hipGetLastError());
}
void mass_mul(mk_ctx* c, int n, const double* mass, const double* u, double* b) {
    hipLaunchKernelGGL(k_mass_mul, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, mass, u, b);
    MK_HIP(hipGetLastError());
}
void mass_dot(mk_ctx* c, int n, const double* mass, const double* x, double* out_dev) {
    const unsigned g = fgrid(size_t(n));
    hipLaunchKernelGGL(k_mass_dot, dim3(g), dim3(256), 0, c->stream, n, mass, x, c->partials.p);
    finish<1, FinOut>(c, g, out_dev);
    MK_HIP(hipGetLastError());
}
void cvt64to32_scaled(mk_ctx* c, size_t len, const double* x, const double* s_dev, float* y) {
    hipLaunchKernelGGL(k_cvt_scaled, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, x, s_dev, y);
    MK_HIP(hipGetLastError());
}
void add32to64_scaled(mk_ctx* c, size_t len, const float* d, const double* s_dev, double* x) {
    hipLaunchKernelGGL(k_add_scaled, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, d, s_dev, x);
    MK_HIP(hipGetLastError());
}

// Generic PCG on the device.  r must hold the initial residual b - A x on entry; bb_dev
// points to ||b||^2 on the device.
static inline void mg_apply(mk_ctx* c, Mg* g, const double* b, double* x, const int* skip) { mg_apply64(c, g, b, x, skip); }
static inline void mg_apply(mk_ctx* c, Mg* g, const float* b, float* x, const int* skip) { mg_apply32(c, g, b, x, skip); }

// q = A p fused with p^H q and alpha (CG state on the device)
template <typename T>
static void launch_cg_spmv(mk_ctx* c, const mk_op* A, const T* val, const T* p, T* q, double bytes) {
    unsigned g = 0;
    const SellDev<T> D = sell_dev<T>(A, val, A->block, &g);
    const unsigned gr = std::max(1u, std::min(g, fin_grid()));
    {
        SpmvTimer tm(c, bytes, A, "cg_spmv", sizeof(T) == 4);
        if (A->cplx)
            MK_TLAUNCH(tm, 0, (cg_spmv<T, true>), dim3(gr), dim3(A->block), 0, c->stream, D, p, q,
                       (const CgState*)c->state.p, c->partials.p);
        else
            MK_TLAUNCH(tm, 0, (cg_spmv<T, false>), dim3(gr), dim3(A->block), 0, c->stream, D, p, q,
                       (const CgState*)c->state.p, c->partials.p);
    }
    finish<1, FinCgAlpha>(c, gr);
}

// Runs `batch` calls of body() on the context stream: eagerly, or -- when graphs are on, timing
// is off and batch is the operator's usual check interval -- by replaying a hipGraph captured
// from the same calls the first time this key is seen.  body() must only launch kernels on
// c->stream whose arguments are fixed by the key: the context (its id and its partials, state
// and scalar buffers), the hierarchy (its id: its level buffers live as long as it does) and
// the solve's vectors and value arrays (`buffers`).  The graphs are cached on the operator, so
// an operator is only ever run on the context it was created on.
template <typename F>
static void run_batch(mk_ctx* c, const mk_op* A_, const Mg* mg, std::vector<uintptr_t> buffers, int batch,
                      bool graphable, F&& body) {
    mk_op* A = const_cast<mk_op*>(A_);
    MK_CHECK(A->ctx == c, MK_ERR_ARG, "operator used with a context other than the one it was created on");
    if (!c->use_graphs || c->time_spmv || !graphable) {
        for (int b = 0; b < batch; b++) body();
        c->eager_batches++;
        return;
    }
    GraphKeyParts kp;
    kp.ctx_uid = c->uid;
    kp.mg_uid = mg_uid(mg);
    kp.state = uintptr_t(c->state.p);
    kp.partials = uintptr_t(c->partials.p);
    kp.scal = uintptr_t(c->scal.p);
    kp.buffers = std::move(buffers);
    kp.batch = batch;
    const std::vector<uintptr_t> key = make_graph_key(kp);
    mk_op::Graph* G = nullptr;
    for (auto& g : A->graphs) if (g.key == key) { G = &g; break; }
    if (!G) {
        if (A->graphs.size() >= 8) {  // bounded cache: drop the oldest
            if (A->graphs.front().exec) MK_HIP(hipGraphExecDestroy(A->graphs.front().exec));
            A->graphs.erase(A->graphs.begin());
        }
        const int64_t calls0 = c->spmv_calls;
        const double bytes0 = c->spmv_bytes;
        const auto t_cap = std::chrono::steady_clock::now();
        MK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        hipGraph_t graph = nullptr;
        try {
            for (int b = 0; b < batch; b++) body();
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        MK_HIP(hipStreamEndCapture(c->stream, &graph));
        mk_op::Graph ng;
        ng.key = key;
        hipError_t e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) throw Error(MK_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        ng.spmv_calls = c->spmv_calls - calls0;
        ng.spmv_bytes = c->spmv_bytes - bytes0;
        c->spmv_calls = calls0;
        c->spmv_bytes = bytes0;
        A->graphs.push_back(std::move(ng));
        G = &A->graphs.back();
        c->graph_captures++;
        MK_LOG("[graph] %s: %d iterations captured and instantiated in %.1f ms\n", A->label.c_str(), batch,
               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_cap).count());
    }
    MK_HIP(hipGraphLaunch(G->exec, c->stream));
    c->spmv_calls += G->spmv_calls;
    c->spmv_bytes += G->spmv_bytes;
    c->graph_replays++;
}

// m iterations as captured batches: the check interval itself, or power-of-two pieces (a few
// graph sizes per operator, each captured once and replayed by every later solve)
template <typename F>
static void run_batches(int m, int interval, F&& batch) {
    if (m == interval) { batch(m); return; }
    for (int p = 1 << 30; m > 0; p >>= 1)
        if (p <= m) { batch(p); m -= p; }
}

// one PCG iteration with an external (multigrid) preconditioner
template <typename T>
static void cg_iter_mg(mk_ctx* c, const mk_op* A, const T* val, Mg* mg, T* x, T* r, T* z, T* p, T* q, double bytes) {
    const int n = A->n;
    const size_t len = size_t(n) * (A->cplx ? 2 : 1);
    const unsigned g = blocks_for(n);
    CgState* st = c->state.p;
    const unsigned gn = fgrid(size_t(n));
    launch_cg_spmv<T>(c, A, val, p, q, bytes);
    if (A->cplx)
        hipLaunchKernelGGL((cg_update_noz<T, true>), dim3(gn), dim3(256), 0, c->stream, n, x, r, p, q,
                           (const CgState*)st, c->partials.p);
    else
        hipLaunchKernelGGL((cg_update_noz<T, false>), dim3(gn), dim3(256), 0, c->stream, n, x, r, p, q,
                           (const CgState*)st, c->partials.p);
    finish<1, FinCgRr>(c, gn);
    mg_apply(c, mg, r, z, &st->done);
    if (A->cplx)
        hipLaunchKernelGGL((cg_beta<T, true>), dim3(gn), dim3(256), 0, c->stream, n, r, z, (const CgState*)st,
                           c->partials.p);
    else
        hipLaunchKernelGGL((cg_beta<T, false>), dim3(gn), dim3(256), 0, c->stream, n, r, z, (const CgState*)st,
                           c->partials.p);
    finish<1, FinCgBeta>(c, gn);
    hipLaunchKernelGGL(cg_pupdate<T>, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, p, z, (const CgState*)st);
}

template <typename T>
static CgResult cg_run(mk_ctx* c, const mk_op* A, const T* val, const T* dinv, T* x, T* r, T* z, T* p, T* q,
                       const double* bb_dev, double rel_tol, int maxiter, int check_every, double bytes,
                       Mg* mg = nullptr) {
    const int n = A->n;
    const size_t len = size_t(n) * (A->cplx ? 2 : 1);
    const unsigned g = blocks_for(n);
    MK_CHECK(g <= kMaxParts, MK_ERR_ARG, "operator too large for the reduction buffer");
    CgState* st = c->state.p;
    hipLaunchKernelGGL(k_set_state_dev, dim3(1), dim3(1), 0, c->stream, st, bb_dev, rel_tol * rel_tol);
    const unsigned gn = fgrid(size_t(n));
    if (mg) {
        mg_apply(c, mg, r, z, nullptr);
        if (A->cplx)
            hipLaunchKernelGGL((cg_init_z<T, true>), dim3(gn), dim3(256), 0, c->stream, n, r, z, p, c->partials.p);
        else
            hipLaunchKernelGGL((cg_init_z<T, false>), dim3(gn), dim3(256), 0, c->stream, n, r, z, p, c->partials.p);
        finish<2, FinCgInitZ>(c, gn);
    } else {
        if (A->cplx)
            hipLaunchKernelGGL((cg_init<T, true>), dim3(gn), dim3(256), 0, c->stream, n, r, dinv, z, p, c->partials.p);
        else
            hipLaunchKernelGGL((cg_init<T, false>), dim3(gn), dim3(256), 0, c->stream, n, r, dinv, z, p, c->partials.p);
        finish<2, FinCgInit>(c, gn);
    }
    MK_HIP(hipGetLastError());
    if (mg) {
        CgState h{};
        BatchPlanner plan(c->time_spmv ? 1 : check_every);
        for (;;) {
            MK_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
            MK_HIP(hipStreamSynchronize(c->stream));
            timing_flush(c);
            if (h.done || h.it >= maxiter) break;
            const int batch = std::min(plan.next(h.it, h.rr, h.tol2 * h.bb), maxiter - h.it);
            run_batches(batch, check_every, [&](int b) {
                run_batch(c, A, mg, {uintptr_t(sizeof(T)), uintptr_t(val), uintptr_t(x), uintptr_t(r),
                                     uintptr_t(z), uintptr_t(p), uintptr_t(q)},
                          b, true, [&] { cg_iter_mg<T>(c, A, val, mg, x, r, z, p, q, bytes); });
            });
            MK_HIP(hipGetLastError());
        }
        CgResult res{h.status, h.it, h.rr, h.bb};
        if (res.status == MK_OK && !h.done) res.status = MK_ERR_NO_CONVERGENCE;
        return res;
    }
    CgState h{};
    int launched = 0;
    BatchPlanner plan(c->time_spmv ? 1 : check_every);
    for (;;) {
        MK_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        MK_HIP(hipStreamSynchronize(c->stream));
        timing_flush(c);
        if (h.done || h.it >= maxiter) break;
        const int batch = std::min(plan.next(h.it, h.rr, h.tol2 * h.bb), maxiter - h.it);
        for (int b = 0; b < batch; b++) {
            launch_cg_spmv<T>(c, A, val, p, q, bytes);
            if (A->cplx)
                hipLaunchKernelGGL((cg_update<T, true>), dim3(gn), dim3(256), 0, c->stream, n, x, r, z, p, q, dinv,
                                   (const CgState*)st, c->partials.p);
            else
                hipLaunchKernelGGL((cg_update<T, false>), dim3(gn), dim3(256), 0, c->stream, n, x, r, z, p, q, dinv,
                                   (const CgState*)st, c->partials.p);
            finish<2, FinCgUpdate>(c, gn);
            hipLaunchKernelGGL(cg_pupdate<T>, dim3(blocks_for(len)), dim3(256), 0, c->stream, len, p, z,
                               (const CgState*)st);
            launched++;
        }
        MK_HIP(hipGetLastError());
    }
    CgResult res{h.status, h.it, h.rr, h.bb};
    if (res.status == MK_OK && !h.done) res.status = MK_ERR_NO_CONVERGENCE;
    return res;
}

CgResult cg64(mk_ctx* c, const mk_op* A, const double* b, double* x, double rel_tol, int maxiter, bool jacobi,
              int check_every, bool x_is_zero, Mg* mg) {
    Work& w = work(c);
    const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
    w.r.ensure(len); w.z.ensure(len); w.p.ensure(len); w.q.ensure(len);
    double* bb = c->scal.p + 0;
    double* rr0 = c->scal.p + 1;
    dot64(c, len, b, b, bb);
    if (x_is_zero) MK_HIP(hipMemcpyAsync(w.r.p, b, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    else spmv64(c, A, x, w.r.p, b, SPMV_RESID, rr0, nullptr);
    return cg_run<double>(c, A, A->val64.p, jacobi ? A->dinv64.p : nullptr, x, w.r.p, w.z.p, w.p.p, w.q.p, bb,
                          rel_tol, maxiter, check_every, A->bytes64, mg);
}

CgResult cg32(mk_ctx* c, const mk_op* A, const float* b, float* x, double rel_tol, int maxiter, bool jacobi,
              int check_every, Mg* mg) {
    Work& w = work(c);
    const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
    w.r32.ensure(len); w.z32.ensure(len); w.p32.ensure(len); w.q32.ensure(len);
    // x = 0, r = b, ||b|| = 1 by construction of the caller (b normalised)
    MK_HIP(hipMemsetAsync(x, 0, len * sizeof(float), c->stream));
    MK_HIP(hipMemcpyAsync(w.r32.p, b, len * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    double one = 1.0;
    MK_HIP(hipMemcpyAsync(c->scal.p + 2, &one, sizeof(double), hipMemcpyHostToDevice, c->stream));
    return cg_run<float>(c, A, A->val32.p, jacobi ? A->dinv32.p : nullptr, x, w.r32.p, w.z32.p, w.p32.p, w.q32.p,
                         c->scal.p + 2, rel_tol, maxiter, check_every, A->bytes32, mg);
}

// FP64 flexible PCG whose multigrid V-cycle runs in FP32 (vectors converted at its ends).
CgResult cg64_mg32(mk_ctx* c, const mk_op* A, const double* b, double* x, double rel_tol, int maxiter,
                   int check_every, Mg* mg) {
    Work& w = work(c);
    const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
    w.r.ensure(len); w.z.ensure(len); w.p.ensure(len); w.q.ensure(len);
    w.r32.ensure(len); w.z32.ensure(len);
    const unsigned g = blocks_for(len);
    MK_CHECK(g <= kMaxParts, MK_ERR_ARG, "operator too large for the reduction buffer");
    double* bb = c->scal.p + 0;
    double* rr0 = c->scal.p + 1;
    dot64(c, len, b, b, bb);
    spmv64(c, A, x, w.r.p, b, SPMV_RESID, rr0, nullptr);
    CgState* st = c->state.p;
    hipLaunchKernelGGL(k_set_state_dev, dim3(1), dim3(1), 0, c->stream, st, (const double*)bb, rel_tol * rel_tol);
    hipLaunchKernelGGL(k_to32, dim3(g), dim3(256), 0, c->stream, len, (const double*)w.r.p, w.r32.p);
    const unsigned gl = fgrid(len);
    hipLaunchKernelGGL(cgm_init2, dim3(gl), dim3(256), 0, c->stream, len, (const double*)w.r.p, w.p.p, w.q.p,
                       w.z.p, st, c->partials.p);
    hipLaunchKernelGGL(cgm_init2_fin, dim3(1), dim3(1024), 0, c->stream, st, (const double*)c->partials.p, int(gl));
    mg_apply32(c, mg, w.r32.p, w.z32.p, &st->done);
    MK_HIP(hipGetLastError());
    unsigned gs = 0;
    const SellDev<double> D = sell_dev<double>(A, A->val64.p, A->block, &gs);
    const unsigned gsr = std::max(1u, std::min(gs, fin_grid()));
    // one iteration: beta, p/q update with A z (fused), x/r update, V-cycle on the new residual
    auto iteration = [&] {
        hipLaunchKernelGGL(cgm_beta2, dim3(gl), dim3(256), 0, c->stream, len, (const double*)w.r.p,
                           (const float*)w.z32.p, (const double*)w.z.p, st, c->partials.p);
        hipLaunchKernelGGL(cgm_beta2_fin, dim3(1), dim3(1024), 0, c->stream, st, (const double*)c->partials.p, int(gl));
        {
            SpmvTimer tm(c, A->bytes64 + double(len) * (4.0 + 8.0 * 5.0), A, "cg_spmv", false);
            if (A->cplx)
                MK_TLAUNCH(tm, 0, (cgm_spmv<true>), dim3(gsr), dim3(A->block), 0, c->stream, D,
                           (const float*)w.z32.p, w.p.p, w.q.p, w.z.p, st, c->partials.p);
            else
                MK_TLAUNCH(tm, 0, (cgm_spmv<false>), dim3(gsr), dim3(A->block), 0, c->stream, D,
                           (const float*)w.z32.p, w.p.p, w.q.p, w.z.p, st, c->partials.p);
        }
        hipLaunchKernelGGL(cgm_spmv_fin, dim3(1), dim3(1024), 0, c->stream, st, (const double*)c->partials.p, int(gsr));
        hipLaunchKernelGGL(cgm_update, dim3(gl), dim3(256), 0, c->stream, len, x, w.r.p, w.r32.p,
                           (const double*)w.p.p, (const double*)w.q.p, st, c->partials.p);
        hipLaunchKernelGGL(cgm_update_fin, dim3(1), dim3(1024), 0, c->stream, st, (const double*)c->partials.p, int(gl));
        mg_apply32(c, mg, w.r32.p, w.z32.p, &st->done);
    };
    CgState h{};
    BatchPlanner plan(c->time_spmv ? 1 : check_every);
    for (;;) {
        MK_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        MK_HIP(hipStreamSynchronize(c->stream));
        timing_flush(c);
        if (h.done || h.it >= maxiter) break;
        const int batch = std::min(plan.next(h.it, h.rr, h.tol2 * h.bb), maxiter - h.it);
        run_batches(batch, check_every, [&](int b) {
            run_batch(c, A, mg, {uintptr_t(4), uintptr_t(A->val64.p), uintptr_t(x), uintptr_t(w.r.p),
                                 uintptr_t(w.r32.p), uintptr_t(w.z32.p), uintptr_t(w.z.p), uintptr_t(w.p.p),
                                 uintptr_t(w.q.p)},
                      b, true, iteration);
        });
        MK_HIP(hipGetLastError());
    }
    CgResult res{h.status, h.it, h.rr, h.bb};
    if (res.status == MK_OK && !h.done) res.status = MK_ERR_NO_CONVERGENCE;
    return res;
}

// Iterative refinement: FP64 residual r = b - A x, FP32 PCG correction on r/||r||.
SolveResult solve_spd(mk_ctx* c, const mk_op* A, const double* b, double* x, const mk_solve_opts& o, Mg* pre,
                      double target) {
    MK_CHECK(o.precision >= MK_PREC_FP64 && o.precision <= MK_PREC_MIXED_MG, MK_ERR_ARG, "unknown precision mode");
    // goal: where the iteration stops; o.tol: what the result must meet (a goal below the FP64
    // floor eps ||A|| ||x|| / ||b|| of the true residual is met as far as that floor allows)
    const double goal = (target > 0.0 && target < o.tol) ? target : o.tol;
    const bool jac = o.precond == MK_PRECOND_JACOBI;
    Mg* mg = (o.precond == MK_PRECOND_AMG) ? (pre ? pre : A->mg) : nullptr;
    MK_CHECK(o.precond != MK_PRECOND_AMG || mg, MK_ERR_ARG, "AMG preconditioner requested but not built");
    SolveResult res{MK_OK, 0, 0, 0.0};
    const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
    // FP32 multigrid inside FP64 PCG; without multigrid there is no FP32 work to do: plain FP64
    const bool mg32 = o.precision == MK_PREC_MIXED_MG && mg;
    if (o.precision == MK_PREC_FP64 || o.precision == MK_PREC_MIXED_MG) {
        // the recurrences can drift from the true residual: it is recomputed in FP64 and the
        // solve restarts from x (at most twice) while it is above the tolerance
        for (int attempt = 0;; attempt++) {
            const CgResult cr = mg32 ? cg64_mg32(c, A, b, x, goal, o.maxiter - res.iters, o.check_every, mg)
                                     : cg64(c, A, b, x, goal, o.maxiter - res.iters, jac, o.check_every, false, mg);
            res.iters += cr.iters;
            Work& w = work(c);
            double h[2];
            w.t.ensure(len);
            spmv64(c, A, x, w.t.p, b, SPMV_RESID, c->scal.p + 3, nullptr);
            MK_HIP(hipMemcpyAsync(h, c->scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            double rr;
            MK_HIP(hipMemcpyAsync(&rr, c->scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            MK_HIP(hipStreamSynchronize(c->stream));
            res.rel_residual = h[0] > 0 ? std::sqrt(rr / h[0]) : std::sqrt(rr);
            res.status = cr.status;
            if (res.status == MK_ERR_NO_CONVERGENCE && res.rel_residual <= o.tol) res.status = MK_OK;
            if (!(res.status == MK_OK && res.rel_residual > o.tol && attempt < 2 && res.iters < o.maxiter))
                break;
            res.refinements++;
        }
        if (res.status == MK_OK && res.rel_residual > o.tol) res.status = MK_ERR_NO_CONVERGENCE;
        return res;
    }
    Work& w = work(c);
    w.t.ensure(len); w.r32.ensure(len); w.x32.ensure(len);
    double* bb = c->scal.p + 4;
    double* rr = c->scal.p + 5;
    dot64(c, len, b, b, bb);
    double hbb = 0.0;
    MK_HIP(hipMemcpyAsync(&hbb, bb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MK_HIP(hipStreamSynchronize(c->stream));
    if (hbb == 0.0) {
        MK_HIP(hipMemsetAsync(x, 0, len * sizeof(double), c->stream));
        MK_HIP(hipStreamSynchronize(c->stream));
        return res;
    }
    double prev = HUGE_VAL;
    int stagn = 0;
    // the FP32 work vectors used by cg32 are r32,z32,p32,q32; the normalised rhs lives in x32's twin
    DevBuf<float>& rhs32 = w.x32;
    DevBuf<float>& d32 = w.d32;
    d32.ensure(len);
    for (;;) {
        spmv64(c, A, x, w.t.p, b, SPMV_RESID, rr, nullptr);
        double hrr = 0.0;
        MK_HIP(hipMemcpyAsync(&hrr, rr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        MK_HIP(hipStreamSynchronize(c->stream));
        const double rel = std::sqrt(hrr / hbb);
        res.rel_residual = rel;
        if (!(rel == rel)) { res.status = MK_ERR_NAN; break; }
        if (rel <= goal) break;
        if (res.refinements >= o.max_refine) { res.status = rel <= o.tol ? MK_OK : MK_ERR_NO_CONVERGENCE; break; }
        if (rel > 0.5 * prev) { if (++stagn >= 3) { res.status = rel <= o.tol ? MK_OK : MK_ERR_STAGNATION; break; } }
        else stagn = 0;
        prev = rel;
        cvt64to32_scaled(c, len, w.t.p, rr, rhs32.p);
        const double itol = std::max(o.inner_tol, 0.5 * goal / rel);
        CgResult cr = cg32(c, A, rhs32.p, d32.p, std::min(itol, 0.5), o.maxiter - res.iters, jac, o.check_every, mg);
        res.iters += cr.iters;
        res.refinements++;
        if (cr.status == MK_ERR_NOT_SPD || cr.status == MK_ERR_NAN) { res.status = cr.status; break; }
        add32to64_scaled(c, len, d32.p, rr, x);
        if (res.iters >= o.maxiter) {
            spmv64(c, A, x, w.t.p, b, SPMV_RESID, rr, nullptr);
            MK_HIP(hipMemcpyAsync(&hrr, rr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            MK_HIP(hipStreamSynchronize(c->stream));
            res.rel_residual = std::sqrt(hrr / hbb);
            if (res.rel_residual > o.tol) res.status = MK_ERR_NO_CONVERGENCE;
            break;
        }
    }
    return res;
}

}  // namespace mk
